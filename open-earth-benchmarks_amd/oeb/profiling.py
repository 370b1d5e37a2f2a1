"""Summaries of rocprofv3 output databases (the SQLite ``*_results.db`` that
rocprofv3 writes) as short markdown for profiles/.

  python -m oeb.profiling stats  <results.db> --command "<cmd>" [--out file.md] [--top N] [--sort avg|total]
          [--skip-first W] [--first F] [--match REGEX] [--dispatches A:B]
      per (kernel, grid size) dispatch count, average / min / max duration, the
      average over each row's dispatches W+1..F, and resources
  python -m oeb.profiling counters <db> [<db> ...] --command "<cmd>" [--match REGEX]
      per (kernel, grid size) mean of every counter of one or more --pmc passes
  python -m oeb.profiling pmc    --fetch <db> --write <db> --copy-bytes N --shape 256x256x80
      --command-fetch "<cmd>" --command-write "<cmd>" [--out-md f.md] [--out-json f.json]
      HBM bytes per launch from FETCH_SIZE / WRITE_SIZE (KiB per dispatch),
      each scaled by what the same counter reads for the plain device copy of
      N bytes that bench.py --copy-mib runs in the same process
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import statistics
import sys
from collections import defaultdict
from pathlib import Path

def short_name(name: str) -> str:
    """'void oeb::(anonymous namespace)::hadvuv_kernel<5>(double*, ...)' -> 'hadvuv_kernel<5>'."""
    s = name.replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    depth, cut = 0, len(s)
    for pos, ch in enumerate(s):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = pos
            break
    s = s[:cut].strip()
    depth, last = 0, 0
    for pos, ch in enumerate(s):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == ":" and depth == 0 and pos + 1 < len(s) and s[pos + 1] == ":":
            last = pos + 2
    return s[last:]


def _rows(db: Path, sql: str):
    con = sqlite3.connect(str(db))
    try:
        return list(con.execute(sql))
    finally:
        con.close()


def kernel_stats(db: Path, dispatches: tuple[int, int] | None = None) -> dict:
    """Dispatches per (kernel, grid size): durations (us, in dispatch order)
    and resources.  One kernel launched on two domain sizes gives two rows.
    `dispatches` = (a, b) keeps only the process's dispatches a..b-1 in start
    order (all kernels counted)."""
    cols = ("name, duration, vgpr_count, accum_vgpr_count, sgpr_count, lds_size, scratch_size, "
            "grid_x, grid_y, grid_z, workgroup_x, workgroup_y, workgroup_z")
    try:
        rows = _rows(db, f"select {cols} from kernels order by start")
    except sqlite3.OperationalError:  # no start column: table order
        rows = _rows(db, f"select {cols} from kernels")
    if dispatches is not None:
        rows = rows[dispatches[0]:dispatches[1]]
    out = defaultdict(lambda: {"durations": []})
    for (name, dur, vgpr, agpr, sgpr, lds, scratch, gx, gy, gz, wx, wy, wz) in rows:
        k = out[(short_name(name), int(gx) * int(gy) * int(gz))]
        k["durations"].append(float(dur) / 1e3)  # ns -> us
        k.update(vgpr=vgpr, agpr=agpr, sgpr=sgpr, lds=lds, scratch=scratch, grid=(gx, gy, gz), wg=(wx, wy, wz))
    return out


def stats_markdown(db: Path, command: str, top: int = 0, skip_first: int = 0, first: int = 0,
                   match: str = "", dispatches: tuple[int, int] | None = None, sort: str = "avg") -> str:
    import re

    st = kernel_stats(db, dispatches)
    if match:
        st = {k: v for k, v in st.items() if re.search(match, k[0])}
    lines = [
        f"# rocprofv3 kernel trace: `{db.name}`",
        "",
        "Command (run from the repository root on the GPU box):",
        "",
        "```",
        command,
        "```",
        "",
        "One row per (kernel, grid size)"
        + (f", of the process's dispatches {dispatches[0]}..{dispatches[1] - 1} in start order" if dispatches else "")
        + (f", the {top} with the largest " + ("total" if sort == "total" else "average") + " time" if top else "")
        + ".  avg / min / max over all dispatches; "
        + (f"'window avg' is the mean of each row's dispatches {skip_first + 1}..{first or 'last'} "
           f"(in dispatch order)." if skip_first or first else ""),
        "",
        "| kernel | grid (threads) | calls | avg us | "
        + ("window avg us | " if skip_first or first else "")
        + "min us | max us | VGPR | SGPR | LDS B | scratch B | workgroup |",
        "|---|---|---|---|" + ("---|" if skip_first or first else "") + "---|---|---|---|---|---|---|",
    ]
    key = (lambda kv: -sum(kv[1]["durations"])) if sort == "total" else (lambda kv: -statistics.mean(kv[1]["durations"]))
    rows = sorted(st.items(), key=key)
    if top:
        rows = rows[:top]
    for (name, grid), k in rows:
        d = k["durations"]
        after = d[skip_first:first or None] or d
        lines.append(
            f"| {name} | {grid} | {len(d)} | {statistics.mean(d):.2f} | "
            + (f"{statistics.mean(after):.2f} | " if skip_first or first else "")
            + f"{min(d):.2f} | {max(d):.2f} | {k['vgpr']} | {k['sgpr']} | {k['lds']} | {k['scratch']} | "
            f"{k['wg'][0]}x{k['wg'][1]} |")
    return "\n".join(lines) + "\n"


def pmc_per_kernel(db: Path, counter: str) -> dict:
    """Average value of `counter` per dispatch, per (kernel, grid size)."""
    rows = _rows(db, "select kernel_name, grid_size, value from counters_collection "
                     f"where counter_name = '{counter}'")
    acc = defaultdict(list)
    for name, grid, val in rows:
        acc[(short_name(name), int(grid))].append(float(val))
    return {k: statistics.mean(v) for k, v in acc.items()}


def counters_markdown(dbs: list, command: str, match: str = "") -> str:
    """Per (kernel, grid size): the mean per dispatch of every counter in the
    rocprofv3 --pmc databases `dbs` (one pass each)."""
    import re

    vals: dict = defaultdict(dict)
    names: list = []
    for db in dbs:
        rows = _rows(Path(db), "select kernel_name, grid_size, counter_name, value from counters_collection")
        acc = defaultdict(list)
        for kname, grid, cname, val in rows:
            acc[(short_name(kname), int(grid), cname)].append(float(val))
        for (k, g, c), v in acc.items():
            if match and not re.search(match, k):
                continue
            vals[(k, g)][c] = statistics.mean(v)
            if c not in names:
                names.append(c)
    lines = [f"# rocprofv3 counters per dispatch", "", "Commands (from the repository root on the GPU box):", "",
             "```", command, "```", "",
             "| kernel | grid (threads) | " + " | ".join(names) + " |", "|---|---|" + "---|" * len(names)]
    for (k, g), cv in sorted(vals.items()):
        lines.append(f"| {k} | {g} | " + " | ".join(f"{cv[c]:.4g}" if c in cv else "" for c in names) + " |")
    return "\n".join(lines) + "\n"


def stencil_of(kernel: str):
    """Which stencil an FP64 forward device kernel belongs to (None for the
    FP32 instantiations, the adjoint kernels and other kernels)."""
    if "float" in kernel or "_adj" in kernel:
        return None
    if kernel.startswith("hadvuv_"):
        return "hadvuv5th" if "<5" in kernel else "hadvuv"
    for name in ("laplace", "hdiffsa", "hdiffsmag", "fastwavesuv", "uvbke", "p_grad_c", "nh_p_grad", "fvtp2d"):
        if kernel.startswith(name + "_"):
            return name
    return None


def main(argv=None) -> int:
    p = argparse.ArgumentParser(prog="python -m oeb.profiling")
    sub = p.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("db")
    s.add_argument("--command", required=True)
    s.add_argument("--out")
    s.add_argument("--top", type=int, default=0, help="only the slowest N rows")
    s.add_argument("--skip-first", type=int, default=0, help="window: leave out each row's first N dispatches")
    s.add_argument("--first", type=int, default=0, help="window: end after each row's Nth dispatch")
    s.add_argument("--match", default="", help="only kernels whose short name matches this regex")
    s.add_argument("--dispatches", default="", help="A:B, only the process's dispatches A..B-1 in start order")
    s.add_argument("--sort", choices=("avg", "total"), default="avg", help="row order (and --top) by average or total time")
    c = sub.add_parser("counters")
    c.add_argument("dbs", nargs="+")
    c.add_argument("--command", required=True)
    c.add_argument("--match", default="")
    c.add_argument("--out")
    m = sub.add_parser("pmc")
    m.add_argument("--fetch", required=True)
    m.add_argument("--write", required=True)
    m.add_argument("--copy-bytes", type=int, required=True)
    m.add_argument("--copy-kernel", default="__amd_rocclr_copyBuffer")
    m.add_argument("--shape", required=True, help="NXxNYxNZ")
    m.add_argument("--command-fetch", required=True)
    m.add_argument("--command-write", required=True)
    m.add_argument("--out-md")
    m.add_argument("--out-json")
    a = p.parse_args(argv)
    if a.cmd == "counters":
        md = counters_markdown(a.dbs, a.command, a.match)
        if a.out:
            Path(a.out).write_text(md)
        print(md)
        return 0
    if a.cmd == "stats":
        disp = tuple(int(x) for x in a.dispatches.split(":")) if a.dispatches else None
        md = stats_markdown(Path(a.db), a.command, a.top, a.skip_first, a.first, a.match, disp, a.sort)
        if a.out:
            Path(a.out).write_text(md)
        print(md)
        return 0
    return _pmc(a)


def _pmc(a) -> int:
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    from oeb.spec import KERNELS

    fetch = pmc_per_kernel(Path(a.fetch), "FETCH_SIZE")
    write = pmc_per_kernel(Path(a.write), "WRITE_SIZE")
    ck = short_name(a.copy_kernel)
    # The calibration copy is the copy-kernel dispatch group that moved the most.
    cands = [k for k in fetch if k[0] == ck and k in write]
    if not cands:
        print(f"copy kernel {ck} not found in the counter runs", file=sys.stderr)
        return 1
    key = max(cands, key=lambda k: write[k])
    # FETCH_SIZE / WRITE_SIZE are KiB per dispatch; the copy reads and writes
    # copy_bytes each, so scale = true bytes / counted bytes.
    f_scale = a.copy_bytes / (fetch[key] * 1024.0)
    w_scale = a.copy_bytes / (write[key] * 1024.0)
    nx, ny, nz = (int(v) for v in a.shape.lower().split("x"))
    pts = nx * ny * nz
    res = {"shape": a.shape, "copy_bytes": a.copy_bytes, "fetch_scale": f_scale, "write_scale": w_scale,
           "command_fetch": a.command_fetch, "command_write": a.command_write, "kernels": {}}
    lines = [f"# HBM traffic per launch, {a.shape} FP64 (rocprofv3 FETCH_SIZE / WRITE_SIZE)", "",
             "Commands (each from the repository root on the GPU box):", "", "```", a.command_fetch,
             a.command_write, "```", "",
             f"Calibration: a device copy of {a.copy_bytes} B read FETCH_SIZE {fetch[key]:.0f} KiB and WRITE_SIZE "
             f"{write[key]:.0f} KiB per dispatch; scale factors {f_scale:.4f} (fetch) and {w_scale:.4f} (write) "
             "convert the counters of every kernel below to bytes.", "",
             "| stencil | device kernels | fetch MB | write MB | HBM MB/launch | algorithmic MB | ratio |",
             "|---|---|---|---|---|---|---|"]
    groups = defaultdict(list)
    for k in fetch:
        st = stencil_of(k[0])
        if st is not None and k in write:
            groups[st].append(k)
    for name in KERNELS:
        ks = groups.get(name)
        if not ks:
            continue
        fb = sum(fetch[k] for k in ks) * 1024.0 * f_scale
        wb = sum(write[k] for k in ks) * 1024.0 * w_scale
        alg = KERNELS[name].bytes_per_point_f64 * pts
        res["kernels"][name] = {"dispatch_kernels": sorted({k[0] for k in ks}), "fetch_bytes": fb, "write_bytes": wb,
                                "hbm_bytes_per_launch": fb + wb, "algorithmic_bytes": alg}
        lines.append(f"| {name} | {', '.join(sorted({k[0] for k in ks}))} | {fb / 1e6:.1f} | {wb / 1e6:.1f} | "
                     f"{(fb + wb) / 1e6:.1f} | {alg / 1e6:.1f} | {(fb + wb) / alg:.3f} |")
    md = "\n".join(lines) + "\n"
    if a.out_md:
        Path(a.out_md).write_text(md)
    if a.out_json:
        Path(a.out_json).write_text(json.dumps(res, indent=1))
    print(md)
    return 0


if __name__ == "__main__":
    sys.exit(main())
