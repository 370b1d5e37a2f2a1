"""The API on the GPU: the reference-signature mirror with the module-level
domain globals, explicit streams, repeatability, and error paths that must
launch nothing."""
import numpy as np
import pytest

from common import copy, inputs, out_region
from gpu_util import to_device
from oeb import abi, stencils
from oeb.layout import Layout
from oeb.spec import KERNELS, ORDER
from oracle import binding as B


@pytest.mark.gpu
def test_reference_signatures_with_globals():
    """laplace(in, out) and hdiffsa(out, in, mask, crlato, crlatu, lap, flx, fly)
    called exactly as the reference drivers call them, extents taken from
    set_domain (the reference's domain_size / domain_height globals)."""
    import torch

    n, nz = 32, 6
    stencils.set_domain(n, nz)
    try:
        for name in ("laplace", "hdiffsa"):
            L, host = inputs(name, n, n, nz, "refdriver", seed=0)
            dev = to_device(name, host, L)
            print(f"[gpu] signature {name}", flush=True)
            if name == "laplace":
                stencils.laplace(dev["in"], dev["out"])
            else:
                stencils.hdiffsa(dev["out"], dev["in"], dev["mask"], dev["crlato"], dev["crlatu"], None, None, None)
            torch.cuda.synchronize()
            ref = copy(host)
            B.run_oracle(name, ref, L)
            o = KERNELS[name].outputs[0]
            assert np.array_equal(out_region(name, o, L, dev[o].to_host()), out_region(name, o, L, ref[o]))
    finally:
        stencils.set_domain(64, 60)


@pytest.mark.gpu
def test_side_stream_and_repeatability():
    """A call on a side stream gives the same doubles as on the default
    stream; two calls in a row give identical outputs."""
    import torch

    name = "fvtp2d"
    L, host = inputs(name, 70, 33, 5, seed=12)
    dev = to_device(name, host, L)
    print("[gpu] side stream fvtp2d", flush=True)
    stencils.run(name, dev, domain=(L.nx, L.ny, L.nz))
    torch.cuda.synchronize()
    first = {o: dev[o].to_host().copy() for o in KERNELS[name].outputs}
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        stencils.run(name, dev, domain=(L.nx, L.ny, L.nz), stream=side.cuda_stream)
    side.synchronize()
    for o in KERNELS[name].outputs:
        a, b = first[o], dev[o].to_host()
        assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]), o


@pytest.mark.gpu
@pytest.mark.parametrize("name", ORDER)
def test_rejected_call_launches_nothing(name):
    """An aliased output is rejected before any launch: the outputs keep their
    NaN fill."""
    import torch

    L, host = inputs(name, 16, 16, 3, seed=1)
    dev = to_device(name, host, L)
    spec = KERNELS[name]
    bad = dict(dev)
    bad[spec.outputs[-1]] = dev[spec.inputs3[0]]
    print(f"[gpu] rejected {name}", flush=True)
    with pytest.raises(abi.OebError) as e:
        stencils.run(name, bad, domain=(16, 16, 3))
    assert e.value.status == abi.OEB_ERR_ALIAS
    torch.cuda.synchronize()
    for o in spec.outputs:
        assert np.isnan(dev[o].to_host()).all()


@pytest.mark.gpu
def test_reference_layout_descriptor_from_util_h():
    """A descriptor laid out exactly as the reference's allocateStorage makes it
    (halo 4 on every axis, k included, no pitch padding) works unchanged."""
    import torch

    from oeb.generators import generate

    name = "nh_p_grad"
    L = Layout(24, 24, 7, h=4, hk=4, align=1)
    assert L.offset == 4 * (1 + L.sj + L.sk) and L.sj == 24 + 8
    host = generate(name, L, "baseline", 2)
    dev = to_device(name, host, L)
    print("[gpu] reference layout nh_p_grad", flush=True)
    stencils.run(name, dev, domain=(24, 24, 7))
    torch.cuda.synchronize()
    ref = copy(host)
    B.run_oracle(name, ref, L)
    for o in KERNELS[name].outputs:
        assert np.array_equal(out_region(name, o, L, dev[o].to_host()), out_region(name, o, L, ref[o]))


@pytest.mark.gpu
@pytest.mark.parametrize("nbytes", [16, 4096 + 48, 37 * 4096, (1 << 24) + 16])
def test_stream_copy(nbytes):
    """oeb_stream_copy moves exactly `nbytes` (partial last workgroup
    included) and leaves the byte after them untouched."""
    import torch

    src = torch.randint(0, 256, (nbytes + 64,), dtype=torch.uint8, device="cuda")
    dst = torch.zeros(nbytes + 64, dtype=torch.uint8, device="cuda")
    print(f"[gpu] stream copy {nbytes} bytes", flush=True)
    abi.check(abi.load().oeb_stream_copy(dst.data_ptr(), src.data_ptr(), nbytes, None))
    torch.cuda.synchronize()
    assert torch.equal(dst[:nbytes], src[:nbytes])
    assert int(dst[nbytes:].count_nonzero()) == 0
