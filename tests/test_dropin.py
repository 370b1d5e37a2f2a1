"""The drop-in C++ layer: the reference's own drivers (src/<name>.cpp) compile
unchanged against this tree's util.h and kernels/<name>.h, also with
ElementType float, and this tree's driver builds.  Compile-only on the CPU;
test_gpu_drivers.py runs them."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "open-earth-benchmarks_amd"
REF = Path("/root/reference")
NAMES = ["laplace", "hdiffsa", "hdiffsmag", "hadvuv", "hadvuv5th", "fastwavesuv", "uvbke", "p_grad_c", "nh_p_grad",
         "fvtp2d"]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not (REF / "src").is_dir(), reason="reference tree not present")
@pytest.mark.parametrize("name", NAMES)
def test_reference_driver_compiles_against_dropin(name, tmp_path):
    """Compiles the reference driver source in place, with this tree's headers
    first on the include path (host code only, -fsyntax-only)."""
    cmd = [HIPCC, "-std=c++17", "-fsyntax-only", f"-I{PKG / 'include'}", f"-I{PKG / 'kernels'}",
           f"-I{ROOT / 'include'}", str(REF / "src" / f"{name}.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.skipif(not (REF / "src").is_dir(), reason="reference tree not present")
@pytest.mark.parametrize("name", NAMES)
def test_reference_driver_compiles_with_float(name, tmp_path):
    """The reference driver with `typedef float ElementType` (the only line
    changed, in a copy under tmp_path) compiles against the drop-in, which
    then calls the FP32 entry points."""
    src = (REF / "src" / f"{name}.cpp").read_text()
    assert "typedef double ElementType;" in src
    f = tmp_path / f"{name}_float.cpp"
    f.write_text(src.replace("typedef double ElementType;", "typedef float ElementType;"))
    cmd = [HIPCC, "-std=c++17", "-fsyntax-only", f"-I{PKG / 'include'}", f"-I{PKG / 'kernels'}",
           f"-I{ROOT / 'include'}", f"-I{REF / 'src'}", str(f)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.parametrize("binary", ["oeb_bench", "oeb_bench_f32"])
def test_own_driver_built(binary):
    assert (PKG / "src" / "oeb_bench.cpp").exists()
    b = PKG / "bin" / binary
    if not b.exists():
        pytest.skip("drivers not built (run __graft_entry__.build())")
    assert b.stat().st_mode & 0o111


def test_driver_table_matches_generator():
    """The driver's fill table (src/oeb_bench.cpp) gives each field the
    fillMath parameters or constant that the oracle's refdriver generator
    uses for it (oeb.generators.REFDRIVER), so their checksums compare."""
    import re

    from oeb.generators import REFDRIVER

    text = (PKG / "src" / "oeb_bench.cpp").read_text()
    sets = {m.group(1): tuple(float(x) for x in m.group(2).split(","))
            for m in re.finditer(r"constexpr Fill (k[A-G])\{Fill::kMath, ([-0-9., ]+)\};", text)}
    assert len(sets) == 7
    for name, fields in REFDRIVER.items():
        block = text[text.index(f't.push_back({{"{name}",'):]
        block = block[:block.index("}});")]
        hadv = re.search(r"hadv_fields\((k[A-G]), (k[A-G])\)", block)
        for field, recipe in fields.items():
            if hadv and field in ("uin", "vin"):
                key, const = hadv.group(1 if field == "uin" else 2), None
            else:
                m = re.search(rf'\{{"{field}", (k[A-G]|value\(([-0-9.]+)\))\}}', block)
                assert m, (name, field)
                key, const = m.group(1), m.group(2)
            if recipe[0] == "fill":
                assert key in sets and sets[key] == tuple(recipe[1:]), (name, field)
            else:
                assert const is not None and float(const) == recipe[1], (name, field)


@pytest.mark.parametrize("name", NAMES)
def test_dropin_header_keeps_reference_signature(name):
    """Every function name of the reference header exists in the drop-in
    header with the same parameter names, in the same order."""
    if not (REF / "kernels").is_dir():
        pytest.skip("reference tree not present")
    import re

    def sigs(text):
        out = {}
        for m in re.finditer(r"void\s+(\w+)\s*\(([^)]*)\)\s*\{", text):
            params = [p.strip().split()[-1].lstrip("&") for p in m.group(2).replace("\n", " ").split(",") if p.strip()]
            out[m.group(1)] = params
        return out

    ref = sigs((REF / "kernels" / f"{name}.h").read_text())
    ours = sigs((PKG / "kernels" / f"{name}.h").read_text())
    ref = {k: v for k, v in ref.items() if k != "advectionDriver"}
    if name == "fvtp2d":
        assert set(ref) == {"fvtp2d", "fvtp2d_fullfusion"}
    for fn, params in ref.items():
        assert fn in ours, fn
        assert ours[fn] == params, fn
