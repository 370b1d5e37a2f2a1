"""The C-ABI library on the CPU: it loads, exports every symbol include/oeb.h
declares, and rejects malformed calls before launching anything.  No GPU is
needed: every case here fails validation, so no kernel is launched.  The
buffers are host (numpy) arrays, handed to the library past the Python
layer's own host-buffer check (check_device=False); the library's last check
rejects host memory, so even a call that passed every other check would
launch nothing (test_host_memory_rejected)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oeb import abi, stencils
from oeb.layout import Layout, Layout1D
from oeb.spec import KERNELS
from oeb.storage import Storage1D, Storage3D

HEADER = Path(__file__).resolve().parent.parent / "include" / "oeb.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oeb_\w+)\s*\(", text)))


def test_library_loads():
    lib = abi.load()
    assert lib.oeb_version().decode().count(".") == 2


def test_every_declared_symbol_is_exported():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 13
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/oeb.h but not exported"


def test_abi_table_matches_header():
    """Every stencil entry in the ctypes table is declared in the header, and
    each parameter list there has the table's number of arrays and scalars."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    for name, sig in abi.SIGNATURES.items():
        m = re.search(rf"int oeb_{name}_f64\((.*?)\);", text, flags=re.S)
        assert m, name
        params = [p.strip() for p in m.group(1).split(",")]
        assert params[0].startswith("const oeb_domain")
        assert params[-1] == "void *stream"
        body = params[1:-1]
        kinds = "".join("3" if "oeb_s3" in p else "1" if "oeb_s1" in p else "d" if p.startswith("double") else "?"
                        for p in body)
        assert kinds == sig, (name, kinds, sig)
        # argument names agree with the registry (the reference's parameter names)
        pnames = [p.split("*")[-1].strip() if "*" in p else p.split()[-1] for p in body]
        arrays = [a.name for a in KERNELS[name].args]
        scalars = [s for s, _ in KERNELS[name].scalars]
        assert pnames == arrays + scalars, name


def test_status_strings():
    lib = abi.load()
    for code in range(8):
        assert lib.oeb_status_string(code)
    assert lib.oeb_status_string(99) == b"unknown status"


def _run(name, f, dom):
    stencils.run(name, f, domain=dom, check_device=False)


def _host(name, nx=8, ny=8, nz=2, **kw):
    L = Layout(nx, ny, nz, **kw)
    L1 = Layout1D(ny, kw.get("h", 3))
    f = {}
    for a in KERNELS[name].args:
        if a.kind == "in1":
            f[a.name] = Storage1D(L1, np.zeros(L1.size))
        elif a.kind in ("in", "out"):
            f[a.name] = Storage3D(L, np.zeros(L.size))
    return L, f


@pytest.mark.parametrize("name", list(KERNELS))
def test_null_array_rejected(name):
    L, f = _host(name)
    first = KERNELS[name].inputs3[0]
    f[first] = None
    with pytest.raises(abi.OebError) as e:
        _run(name, f, (8, 8, 2))
    assert e.value.status == abi.OEB_ERR_NULL and first in e.value.message


@pytest.mark.parametrize("name", list(KERNELS))
def test_output_aliasing_input_rejected(name):
    L, f = _host(name)
    f[KERNELS[name].outputs[0]] = f[KERNELS[name].inputs3[0]]
    with pytest.raises(abi.OebError) as e:
        _run(name, f, (8, 8, 2))
    assert e.value.status == abi.OEB_ERR_ALIAS


@pytest.mark.parametrize("name", list(KERNELS))
def test_too_small_halo_rejected(name):
    """A layout without the halo the stencil reads is rejected (a k-halo of 0
    for the k-reaching stencils, an i/j halo below the reach otherwise)."""
    spec = KERNELS[name]
    if spec.hk:
        L, f = _host(name, hk=0)
    else:
        L, f = _host(name, h=spec.h - 1, align=1)
    with pytest.raises(abi.OebError) as e:
        _run(name, f, (8, 8, 2))
    assert e.value.status == abi.OEB_ERR_LAYOUT


def test_domain_rejected():
    L, f = _host("laplace")
    for dom in [(0, 8, 2), (8, -1, 2), (8, 8, 0)]:
        with pytest.raises(abi.OebError) as e:
            _run("laplace", f, dom)
        assert e.value.status == abi.OEB_ERR_DOMAIN
    with pytest.raises(abi.OebError) as e:  # more rows than the allocation holds
        _run("laplace", f, (8, 12, 2))
    assert e.value.status == abi.OEB_ERR_LAYOUT


def test_mismatched_strides_rejected():
    L, f = _host("uvbke")
    L2 = Layout(8, 8, 2, align=1)
    f["cosa"] = Storage3D(L2, np.zeros(L2.size))
    with pytest.raises(abi.OebError) as e:
        _run("uvbke", f, (8, 8, 2))
    assert e.value.status == abi.OEB_ERR_LAYOUT and "strides" in e.value.message


def test_bad_descriptor_fields_rejected():
    L, f = _host("laplace")
    d = f["in"].desc
    d.strides[2] = 2
    with pytest.raises(abi.OebError) as e:
        _run("laplace", f, (8, 8, 2))
    assert e.value.status == abi.OEB_ERR_LAYOUT
    d.strides[2] = 1
    d.offset = -1
    with pytest.raises(abi.OebError):
        _run("laplace", f, (8, 8, 2))


def test_one_dimensional_too_short_rejected():
    L, f = _host("hdiffsa")
    short = Layout1D(8, 0)
    f["crlato"] = Storage1D(short, np.zeros(short.size))
    with pytest.raises(abi.OebError) as e:
        _run("hdiffsa", f, (8, 8, 2))
    assert e.value.status == abi.OEB_ERR_LAYOUT and "crlato" in e.value.message


@pytest.mark.parametrize("name", list(KERNELS))
def test_host_memory_rejected(name):
    """Well-formed descriptors of host (pageable) memory: the library rejects
    the call with OEB_ERR_MEMORY instead of launching a kernel that would
    fault on them; the Python layer rejects them earlier with TypeError."""
    L, f = _host(name)
    with pytest.raises(abi.OebError) as e:
        _run(name, f, (8, 8, 2))
    assert e.value.status == abi.OEB_ERR_MEMORY
    with pytest.raises(TypeError, match="host buffer"):
        stencils.run(name, f, domain=(8, 8, 2))


def test_last_error_is_per_call():
    lib = abi.load()
    L, f = _host("laplace")
    f["out"] = f["in"]
    with pytest.raises(abi.OebError):
        _run("laplace", f, (8, 8, 2))
    assert b"overlaps" in lib.oeb_last_error()


def test_reference_signature_mirror():
    """The Python mirror keeps the reference's parameter names and order."""
    import inspect

    for name, spec in KERNELS.items():
        params = [p for p in inspect.signature(getattr(stencils, name)).parameters
                  if p not in ("domain", "stream")]
        want = [a.name if a.name != "in" else "in_" for a in spec.args] + [s for s, _ in spec.scalars]
        assert params == want, name


def test_stream_copy_validation():
    """oeb_stream_copy rejects NULL pointers and sizes or addresses that are
    not multiples of 16 before touching memory; a zero-byte copy is a no-op."""
    lib = abi.load()
    a = np.zeros(64, dtype=np.float64)
    b = np.zeros(64, dtype=np.float64)
    pa, pb = a.ctypes.data, b.ctypes.data
    assert lib.oeb_stream_copy(None, pb, 64, None) == abi.OEB_ERR_NULL
    assert lib.oeb_stream_copy(pa, None, 64, None) == abi.OEB_ERR_NULL
    assert lib.oeb_stream_copy(pa, pb, 24, None) == abi.OEB_ERR_ARG
    assert lib.oeb_stream_copy(pa + 8, pb, 64, None) == abi.OEB_ERR_ARG
    assert lib.oeb_stream_copy(pa, pb, 0, None) == abi.OEB_OK


def test_stream_copy_rejects_overlap_and_host_memory():
    lib = abi.load()
    a = np.zeros(64, dtype=np.float64)
    pa = a.ctypes.data
    assert lib.oeb_stream_copy(pa, pa + 128, 256, None) == abi.OEB_ERR_ALIAS
    b = np.zeros(64, dtype=np.float64)
    assert lib.oeb_stream_copy(pa, b.ctypes.data, 256, None) == abi.OEB_ERR_MEMORY


def test_f32_and_f32d_tables_match_header():
    """The FP32 and deviation-island entry points: each parameter's type in
    the header matches the ctypes table ('3' oeb_s3f, 'D' oeb_s3, 'E' oeb_s3f,
    'b' oeb_s1, '1' oeb_s1f, 'f' float, 'd' double)."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)

    def kind(p):
        for pat, k in (("oeb_s3f", "3"), ("oeb_s3 ", "D"), ("oeb_s1f", "1"), ("oeb_s1 ", "b")):
            if pat in p + " ":
                return k
        return "f" if p.startswith("float") else "d" if p.startswith("double") else "?"

    for suffix, table in (("f32", abi.SIGNATURES_F32), ("f32d", abi.SIGNATURES_F32D)):
        for name, sig in table.items():
            m = re.search(rf"int oeb_{name}_{suffix}\((.*?)\);", text, flags=re.S)
            assert m, (name, suffix)
            params = [p.strip() for p in m.group(1).split(",")][1:-1]
            kinds = "".join(kind(p) for p in params)
            want = sig.replace("E", "3")
            assert kinds == want, (name, suffix, kinds, want)
            if suffix == "f32d":
                # each island is followed by its profile, named <island>0
                pnames = [p.split("*")[-1].strip() for p in params if "*" in p]
                islands = [n for n, k in zip(pnames, kinds) if k == "b"]
                assert islands and all(n.endswith("0") and n[:-1] in pnames for n in islands), (name, islands)


def test_f32d_validation():
    """The deviation entry points reject a NULL or too-short base profile
    before launching anything (host buffers: no GPU needed)."""
    import torch  # noqa: F401  (the library binds to torch's HIP runtime)

    from oeb.generators import cast, generate
    from oeb.islands import split_host
    from oeb.storage import DeviationIsland, level_layout

    L = Layout(16, 12, 4, h=3, hk=1)
    host = cast("p_grad_c", generate("p_grad_c", L, "baseline", 0), "f32")
    arrays = {}
    for a in KERNELS["p_grad_c"].args:
        if a.kind in ("in", "out"):
            if a.name in ("gz", "pkc"):
                dev, base = split_host(L, host[a.name])
                arrays[a.name] = DeviationIsland(Storage3D(L, dev), Storage1D(level_layout(L), base))
            else:
                arrays[a.name] = Storage3D(L, host[a.name])
    with pytest.raises(abi.OebError) as e:
        stencils.run("p_grad_c", arrays, domain=(16, 12, 4), check_device=False)
    assert e.value.status == abi.OEB_ERR_MEMORY  # host memory, rejected last
    short = Layout1D(3, 1)  # covers levels [-1, 4): level nz = 4 is missing
    arrays["gz"] = DeviationIsland.__new__(DeviationIsland)
    arrays["gz"].dev = Storage3D(L, split_host(L, host["gz"])[0])
    arrays["gz"].base = Storage1D(short, np.zeros(short.size))
    with pytest.raises(abi.OebError) as e:
        stencils.run("p_grad_c", arrays, domain=(16, 12, 4), check_device=False)
    assert e.value.status == abi.OEB_ERR_LAYOUT and "gz0" in str(e.value)
    with pytest.raises(ValueError):
        DeviationIsland(arrays["gz"].dev, Storage1D(short, np.zeros(short.size)))
    mixed = dict(arrays)
    mixed["gz"] = Storage3D(L, host["gz"])
    with pytest.raises(TypeError, match="all FP64 arrays or all DeviationIsland"):
        stencils.run("p_grad_c", mixed, domain=(16, 12, 4), check_device=False)
