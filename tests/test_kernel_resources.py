"""Register budget of the built library, read from its own code objects.

liboeb.so carries one offload bundle per source file in its .hip_fatbin
section; each bundle's gfx950 code object holds an AMDGPU metadata note
with every kernel's register counts.  No kernel may spill vector registers
or use per-lane scratch, except the forms in SPILLING, which no stencil
takes by default.

A spill moves registers through scratch memory on every use.  fvtp2d's
hoisted tile runs at its 6-waves-per-SIMD cap with 78 of 80 VGPRs; one
more branch in the shared tile-coordinate code made it spill 6 VGPRs and
run 15% slower (596 against 519 us at 512x512x80), with no other sign."""
import re
import shutil
import struct
import subprocess
from pathlib import Path

import pytest
import yaml

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "open-earth-benchmarks_amd" / "lib" / "liboeb.so"
LLVM = Path("/opt/rocm/llvm/bin")
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"

# Forms that spill by design, each taken only through a knob: the adjoint row
# marches with rows loaded two ahead (depth 2), and the FP64 hadvuv5th
# adjoint's row march (FP64 takes the tiles by default).
SPILLING = {
    "hadvuv5th_adj_strip<double, 1>",
    "hadvuv5th_adj_strip<double, 2>",
    "hdiffsmag_adj_strip<double, 2>",
}


def _code_objects(fatbin: bytes):
    """The gfx950 code objects of every offload bundle in a .hip_fatbin."""
    for m in re.finditer(re.escape(MAGIC), fatbin):
        s = m.start()
        (n,) = struct.unpack_from("<Q", fatbin, s + len(MAGIC))
        p = s + len(MAGIC) + 8
        for _ in range(n):
            off, size, tlen = struct.unpack_from("<QQQ", fatbin, p)
            p += 24
            triple = fatbin[p:p + tlen].decode()
            p += tlen
            if triple.endswith("gfx950") and size:
                yield fatbin[s + off:s + off + size]


def _kernels(tmp_path):
    """(demangled name, metadata) of every kernel in liboeb.so."""
    for tool in ("llvm-objcopy", "llvm-readelf"):
        if not (LLVM / tool).exists():
            pytest.skip(f"{tool} not found under {LLVM}")
    if not LIB.exists():
        pytest.skip("liboeb.so is not built")
    fb = tmp_path / "fatbin"
    subprocess.run([str(LLVM / "llvm-objcopy"), f"--dump-section=.hip_fatbin={fb}", str(LIB), str(tmp_path / "so")],
                   check=True, capture_output=True)
    out = []
    for i, co in enumerate(_code_objects(fb.read_bytes())):
        f = tmp_path / f"co{i}.elf"
        f.write_bytes(co)
        notes = subprocess.run([str(LLVM / "llvm-readelf"), "--notes", str(f)], check=True, capture_output=True,
                               text=True).stdout
        doc = notes[notes.index("---"):notes.index("\n...", notes.index("---"))]
        out += yaml.safe_load(doc)["amdhsa.kernels"]
    names = [k[".name"] for k in out]
    filt = shutil.which("c++filt")
    if filt:
        names = subprocess.run([filt], input="\n".join(names), capture_output=True, text=True,
                               check=True).stdout.splitlines()
    return list(zip(names, out))


def _short(name: str) -> str:
    """`void oeb::(anonymous namespace)::f<double, 1>(args)` -> `f<double, 1>`."""
    name = name.removeprefix("void ").replace("oeb::(anonymous namespace)::", "").replace("oeb::", "")
    depth = 0
    for i, ch in enumerate(name):
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0:
            return name[:i]
    return name


def test_no_kernel_spills(tmp_path):
    kernels = _kernels(tmp_path)
    # Every source file's kernels are there (ten stencils, their FP32 and
    # batched forms, adjoints, halo and copy kernels).
    assert len(kernels) > 150, f"only {len(kernels)} kernels parsed from {LIB}"
    bad = []
    for name, md in kernels:
        short = _short(name)
        if short in SPILLING:
            continue
        if md[".vgpr_spill_count"] or md[".private_segment_fixed_size"]:
            bad.append(f"{short}: {md['.vgpr_spill_count']} VGPRs spilled, "
                       f"{md['.private_segment_fixed_size']} B scratch per lane, {md['.vgpr_count']} VGPRs")
    assert not bad, "kernels that spill:\n" + "\n".join(bad)


def test_spilling_forms_are_present(tmp_path):
    """The allowed spillers still exist under these names (a rename would
    otherwise leave SPILLING stale and silently widen it)."""
    short = {_short(n) for n, _ in _kernels(tmp_path)}
    assert SPILLING <= short, f"not in the library: {sorted(SPILLING - short)}"
