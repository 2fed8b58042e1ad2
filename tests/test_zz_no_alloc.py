"""mdmc_backward_noalloc against mdmc_backward through the C ABI.

The no-allocation entry point keeps its scratch in the library's fixed per-device array;
its result must equal mdmc_backward's bit for bit.  Two extractions in a row, the second on
a smaller grid, so the second call runs over scratch the first one left behind; with and
without padding (padding adds the processed-grid gradient to the scratch), overwriting and
accumulating.  A problem past the scratch size is refused with MDMC_ERR_CAPACITY.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAPACITY = 3


def run(d, field, iso, pad, seed, accumulate):
    import torch

    import manifolddmc_amd as mdmc
    nz, ny, nx = field.shape
    n = max(field.shape)
    flags = mdmc.RECORD_ADJOINT | (mdmc.PAD1 if pad else 0)
    d.extract_raw(field.data_ptr(), (nx, ny, nz), (0.0, 0.0, 0.0), (1.0 / (n - 1),) * 3, iso, flags, 0, True)
    V = d.info()[0]
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    dpos = torch.randn((V, 3), device="cuda", generator=g)
    start = torch.randn(field.shape, device="cuda", generator=g)
    a = start.clone() if accumulate else torch.full(field.shape, float("nan"), device="cuda")
    b = a.clone()
    torch.cuda.synchronize()
    d.backward(dpos.data_ptr(), a.data_ptr(), accumulate)
    d.backward_noalloc(dpos.data_ptr(), b.data_ptr(), accumulate)
    torch.cuda.synchronize()
    return V, a, b


@pytest.mark.parametrize("pad", [True, False])
@pytest.mark.parametrize("accumulate", [False, True])
def test_noalloc_backward_equals_backward(pad, accumulate):
    import torch

    import manifolddmc_amd as mdmc
    from manifolddmc_amd import grids
    d = mdmc.DualMarchingCubes()
    try:
        for k, n in enumerate((96, 40)):
            print("no-alloc backward: spheres %d^3, pad %s, accumulate %s" % (n, pad, accumulate), flush=True)
            field = torch.from_numpy(grids.sphere_sdf(n, n_spheres=12, seed=7 + k)).cuda()
            V, a, b = run(d, field, 0.0, pad, 100 + k, accumulate)
            assert V > 0
            assert torch.isfinite(a).all()
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (n, (a - b).abs().max().item())
    finally:
        d.close()


def test_noalloc_backward_refuses_past_its_scratch():
    import torch

    import manifolddmc_amd as mdmc
    n = 330  # padded: 332^3 = 36.6 M grid points, more than the 32 Mi floats of scratch
    field = torch.empty((n, n, n), device="cuda").fill_(1.0)
    field[n // 2, n // 2, n // 2] = -1.0
    d = mdmc.DualMarchingCubes()
    try:
        d.extract_raw(field.data_ptr(), (n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, mdmc.RECORD_ADJOINT | mdmc.PAD1, 0, True)
        V = d.info()[0]
        dpos = torch.zeros((max(V, 1), 3), device="cuda")
        out = torch.zeros_like(field)
        torch.cuda.synchronize()
        with pytest.raises(mdmc.MdmcError) as e:
            d.backward_noalloc(dpos.data_ptr(), out.data_ptr(), False)
        assert e.value.status == CAPACITY
        d.backward(dpos.data_ptr(), out.data_ptr(), False)  # the allocating entry point still serves it
        torch.cuda.synchronize()
    finally:
        d.close()
