"""Grid-pass tile shapes and the autotuner: every shape gives the oracle's mesh, and the
autotuned context gives the same arrays as a fixed shape."""
import numpy as np
import pytest

import meshcheck
import oracle_api as oa

pytestmark = pytest.mark.gpu

ARRAYS = ("positions", "normals", "halfedges", "face_halfedge", "face_attr", "vertex_flags", "vertex_halfedge")


def fields():
    from manifolddmc_amd import grids
    rng = np.random.default_rng(11)
    out = [("random 9x10x11", rng.uniform(-1, 1, size=(9, 10, 11)).astype(np.float32), 0.0)]
    # several tiles of every shape, a partial last tile, rows and slices that straddle tiles
    g = grids.gyroid(45, periods=3, noise=0.05, seed=3)[:37, :45, :41]
    out.append(("gyroid 41x45x37", np.ascontiguousarray(g, np.float32), 0.0))
    b = grids.gaussian_blobs(40, n_blobs=6, seed=4)[:, :33, :]
    out.append(("blobs 40x33x40", np.ascontiguousarray(b, np.float32), 0.5))
    # rows longer than a row-slab band (254 cells): 3 bands, halo cells between them
    w = grids.gyroid(600, periods=40, noise=0.05, seed=5, z0=0, z1=6)[:, :11, :]
    out.append(("gyroid 600x11x6", np.ascontiguousarray(w, np.float32), 0.0))
    return out


def extract(d, f, iso, flags):
    nz, ny, nx = f.shape
    st = d.extract_raw(f.ctypes.data, (nx, ny, nz), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), iso, flags, 0, False)
    return d.mesh_numpy(), st.as_dict()


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_every_tile_shape_matches_oracle(shape):
    import manifolddmc_amd as mdmc
    d = mdmc.DualMarchingCubes()
    try:
        d.set_tile_shape(shape)
        for flags in (oa.PAD1 | oa.SURF_ITER | oa.ALL_POST, oa.SURF_ITER | oa.SPLIT_NM):
            for name, f, iso in fields():
                print("shape", shape, mdmc.TILE_SHAPES[shape], "case", name, "flags", flags, flush=True)
                g, st = extract(d, f, iso, flags)
                assert d.tile_shape()[0] == shape
                o = oa.extract(f, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), iso, flags=flags)
                errs = [e for e in (meshcheck.exact_diff(n, a, b) for n, a, b in
                                    [("positions", g.positions, o.pos), ("normals", g.normals, o.nrm), ("halfedges", g.halfedges, o.he),
                                     ("face_halfedge", g.face_halfedge, o.face_he), ("face_attr", g.face_attr, o.face_attr),
                                     ("vertex_flags", g.vertex_flags, o.vflags), ("vertex_halfedge", g.vertex_halfedge, o.vhe)]) if e]
                assert st["n_dmc_vertices"] == o.stats["n_dmc_vertices"] and st["n_dmc_quads"] == o.stats["n_dmc_quads"]
                assert not errs, (name, errs)
    finally:
        d.close()


def test_autotuner_picks_a_shape_once_per_class_and_keeps_the_mesh():
    import manifolddmc_amd as mdmc
    from manifolddmc_amd import grids
    f = np.ascontiguousarray(grids.gyroid(64, periods=4, noise=0.05, seed=9), np.float32)
    flags = oa.PAD1 | oa.SURF_ITER | oa.ALL_POST
    auto = mdmc.DualMarchingCubes()
    fixed = mdmc.DualMarchingCubes()
    try:
        print("autotune gyroid 64^3", flush=True)
        g1, _ = extract(auto, f, 0.0, flags)
        shape, ms = auto.tile_shape()
        assert shape in mdmc.TILE_SHAPES
        assert all(v > 0 for v in ms.values()), ms
        # the fastest shape, a launch-order or work-queue tile shape (4-6, 9) counted 3% slower
        # than its time
        cost = {k: v * (1.03 if 4 <= k <= 6 or k == 9 else 1.0) for k, v in ms.items()}
        assert cost[shape] == min(cost.values())
        g2, _ = extract(auto, f, 0.1, flags)  # same class: no new tuning
        assert auto.tile_shape() == (shape, ms)
        fixed.set_tile_shape(2 if shape != 2 else 5)
        h1, _ = extract(fixed, f, 0.0, flags)
        h2, _ = extract(fixed, f, 0.1, flags)
        for name in ARRAYS:
            assert meshcheck.exact_diff(name, getattr(g1, name), getattr(h1, name)) is None
            assert meshcheck.exact_diff(name, getattr(g2, name), getattr(h2, name)) is None
        with pytest.raises(mdmc.MdmcError):
            fixed.set_tile_shape(len(mdmc.TILE_SHAPES) + 1)
    finally:
        auto.close()
        fixed.close()


@pytest.mark.parametrize("shape", [1, 3, 5, 7, 8, 9])
def test_batch_with_fixed_shape_equals_single_extractions(shape):
    import manifolddmc_amd as mdmc
    from manifolddmc_amd import grids
    rng = np.random.default_rng(21)
    fs = []
    for i in range(5):
        nz, ny, nx = (int(v) for v in rng.integers(5, 40, size=3))
        f = grids.gyroid(max(nx, ny, nz), periods=2 + i, noise=0.05, seed=i)[:nz, :ny, :nx]
        fs.append(np.ascontiguousarray(f, np.float32))
    isos = [[-0.2, 0.0, 0.3]] * len(fs)
    flags = oa.PAD1 | oa.SURF_ITER | oa.ALL_POST
    b = mdmc.DualMarchingCubes()
    one = mdmc.DualMarchingCubes()
    try:
        b.set_tile_shape(shape)
        print("batch shape", shape, [f.shape for f in fs], flush=True)
        b.extract_batch(fs, isos, [(0, 0, 0)] * len(fs), [(1, 1, 1)] * len(fs), flags)
        meshes = b.batch_meshes()
        k = 0
        for f in fs:
            for iso in isos[0]:
                m, _ = extract(one, f, iso, flags)
                for name in ARRAYS:
                    e = meshcheck.exact_diff(name, getattr(meshes[k], name), getattr(m, name))
                    assert e is None, (k, e)
                k += 1
    finally:
        b.close()
        one.close()
