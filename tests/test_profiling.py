"""oeb.profiling on small synthetic rocprofv3 databases (CPU): kernel-name
shortening, per (kernel, grid) statistics, the per-row window, the global
dispatch window, sorting, and the counter merge."""
import sqlite3

from oeb import profiling as P


def _trace_db(path, rows):
    con = sqlite3.connect(str(path))
    con.execute("create table kernels (name text, start integer, duration integer, vgpr_count integer, "
                "accum_vgpr_count integer, sgpr_count integer, lds_size integer, scratch_size integer, "
                "grid_x integer, grid_y integer, grid_z integer, workgroup_x integer, workgroup_y integer, "
                "workgroup_z integer)")
    for name, start, dur, grid in rows:
        con.execute("insert into kernels values (?,?,?,40,0,48,0,0,?,1,1,256,1,1)", (name, start, dur, grid))
    con.commit()
    con.close()
    return path


def test_short_name():
    assert P.short_name("void oeb::(anonymous namespace)::hadvuv_kernel<5, double, 1024>(oeb::HadvArgs<double>, "
                        "oeb::Geo)") == "hadvuv_kernel<5, double, 1024>"
    assert P.short_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_stats_windows_and_sort(tmp_path):
    # A: 5 dispatches of 10, 20, 30, 40, 50 us; B: 2 of 100 us; C after them.
    rows = [("void oeb::a_kernel<double>(int)", t, 1000 * (10 * (t + 1)), 64) for t in range(5)]
    rows += [("void oeb::b_kernel<double>(int)", 10 + t, 100000, 128) for t in range(2)]
    rows += [("void oeb::c_kernel<float>(int)", 20, 1000, 64)]
    db = _trace_db(tmp_path / "run_results.db", rows)
    st = P.kernel_stats(db)
    assert st[("a_kernel<double>", 64)]["durations"] == [10.0, 20.0, 30.0, 40.0, 50.0]
    md = P.stats_markdown(db, "cmd", skip_first=1, first=4)
    line = next(ln for ln in md.splitlines() if ln.startswith("| a_kernel"))
    assert "| 5 | 30.00 | 30.00 |" in line  # avg of all 5, window 2..4 = (20+30+40)/3
    # global window: the first 6 dispatches in start order are A x5 and one B
    st6 = P.kernel_stats(db, (0, 6))
    assert len(st6[("b_kernel<double>", 128)]["durations"]) == 1 and ("c_kernel<float>", 64) not in st6
    # sorted by total time: A totals 150 us, B 200 us, C 1 us
    md_t = P.stats_markdown(db, "cmd", top=2, sort="total")
    names = [ln.split("|")[1].strip() for ln in md_t.splitlines() if ln.startswith("| ") and "kernel |" not in ln]
    assert names == ["b_kernel<double>", "a_kernel<double>"]


def test_counters_merge(tmp_path):
    dbs = []
    for n, (cname, vals) in enumerate((("SQ_WAVES", (8, 12)), ("SQ_INSTS_VALU", (100, 300)))):
        p = tmp_path / f"pmc{n}.db"
        con = sqlite3.connect(str(p))
        con.execute("create table counters_collection (kernel_name text, grid_size integer, counter_name text, "
                    "value real)")
        for v in vals:
            con.execute("insert into counters_collection values (?,?,?,?)", ("void oeb::k_kernel<float>(int)",
                                                                           256, cname, v))
        con.execute("insert into counters_collection values (?,?,?,?)", ("other", 1, cname, 1))
        con.commit()
        con.close()
        dbs.append(p)
    md = P.counters_markdown(dbs, "cmd", match="^k_kernel")
    line = next(ln for ln in md.splitlines() if ln.startswith("| k_kernel"))
    assert line == "| k_kernel<float> | 256 | 10 | 200 |"
    assert "other" not in md
