"""The processed grid (UniformGridHost::readFromFile + refineGrid, UniformGrid.inl:573-608,
:279-319) read back with mdmc_processed_grid, bit for bit against the oracle's.

Without refinement the padding pass copies each point's first corner and falls back to the
full zero-weight trilinear blend when the input holds a value that is not "plain" (|v| >= 2^126,
-0, inf or NaN); the special-value cases below take that fallback and must still give the
oracle's bits (the blend turns some -0 into +0).
"""
import numpy as np
import pytest

import oracle_api as oa

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def random_field(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, size=shape).astype(np.float32)


def with_specials(v, seed):
    """-0 at random points and on the border, and two finite values beyond 2^126."""
    v = v.copy()
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, v.size, size=max(4, v.size // 50))
    v.reshape(-1)[idx] = np.float32(-0.0)
    v[0, 0, 0] = np.float32(-0.0)
    v[-1, -1, -1] = np.float32(-0.0)
    v.reshape(-1)[v.size // 3] = np.float32(1.5e38)
    v.reshape(-1)[v.size // 2] = np.float32(-1.5e38)
    return v


def gpu_grid(v, origin, delta, iso, flags, refinement):
    import manifolddmc_amd as mdmc
    d = mdmc.DualMarchingCubes()
    try:
        d.extract_raw(v.ctypes.data, (v.shape[2], v.shape[1], v.shape[0]), origin, delta, iso, flags, refinement, False)
        return d.processed_grid()
    finally:
        d.close()


CASES = [(True, 0), (False, 0), (True, 1), (False, 2)]


@pytest.mark.parametrize("special", [False, True])
@pytest.mark.parametrize("pad,refinement", CASES)
def test_processed_grid_equals_oracle(pad, refinement, special):
    import manifolddmc_amd as mdmc
    v = random_field((9, 11, 13), 5)
    if special:
        v = with_specials(v, 6)
    origin, delta = (0.5, -1.0, 2.0), (0.25, 0.5, 1.0)
    flags = mdmc.PAD1 if pad else 0
    vals, o, dl = gpu_grid(v, origin, delta, 0.0, flags, refinement)
    ref = oa.extract(v, origin, delta, 0.0, flags=oa.PAD1 if pad else 0, refinement=refinement, want_grid=True)
    assert vals.shape == ref.grid.shape
    assert np.array_equal(bits(vals), bits(ref.grid))
    assert np.array_equal(np.float32(o), ref.grid_origin) and np.array_equal(np.float32(dl), ref.grid_delta)
    if refinement == 0:
        # a plain copy of the first corners equals the blend exactly when no value is special;
        # with -0 in the input the blend turns some of them into +0, so the copy alone would fail
        copy = np.pad(v, 1, mode="edge") if pad else v
        assert np.array_equal(bits(copy), bits(ref.grid)) == (not special)


def test_special_values_mesh_equals_oracle():
    # the whole pipeline on a field that takes the exact padding fallback
    import manifolddmc_amd as mdmc
    from test_gpu_parity import assert_same
    v = with_specials(random_field((14, 12, 10), 8), 9)
    assert_same(v, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST, label="specials")


def test_batch_processed_grids():
    import manifolddmc_amd as mdmc
    fields = [random_field((7, 8, 9), 1), with_specials(random_field((10, 6, 5), 2), 3)]
    origins = [(0.0, 0.0, 0.0), (1.0, 2.0, 3.0)]
    deltas = [(1.0, 1.0, 1.0), (0.5, 0.25, 2.0)]
    d = mdmc.DualMarchingCubes()
    try:
        d.extract_batch(fields, [[0.0, 0.25], [0.1, -0.2]], origins, deltas, mdmc.PAD1 | mdmc.SURF_ITER)
        for i, v in enumerate(fields):
            vals, o, dl = d.processed_grid(i)
            ref = oa.extract(v, origins[i], deltas[i], 0.0, flags=oa.PAD1, want_grid=True)
            assert np.array_equal(bits(vals), bits(ref.grid)), i
            assert np.array_equal(np.float32(o), ref.grid_origin) and np.array_equal(np.float32(dl), ref.grid_delta)
        with pytest.raises(mdmc.MdmcError):
            d.processed_grid(2)
    finally:
        d.close()
