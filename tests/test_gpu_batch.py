"""The batched form against single extractions: every problem of mdmc_extract_batch
must come out exactly as mdmc_extract gives it (and single extractions equal the
oracle, see test_gpu_parity.py)."""
import numpy as np
import pytest

import meshcheck
import oracle_api as oa

pytestmark = pytest.mark.gpu


def mixed_fields(seed, count):
    from manifolddmc_amd import grids
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nz, ny, nx = (int(v) for v in rng.integers(6, 28, size=3))
        kind = i % 3
        if kind == 0:
            f = grids.gaussian_blobs(max(nx, ny, nz), n_blobs=int(rng.integers(2, 6)), seed=int(rng.integers(1 << 30)))[:nz, :ny, :nx]
        elif kind == 1:
            f = grids.gyroid(max(nx, ny, nz), periods=int(rng.integers(1, 4)), noise=0.05, seed=int(rng.integers(1 << 30)))[:nz, :ny, :nx]
        else:
            f = rng.uniform(-1, 1, size=(nz, ny, nx)).astype(np.float32)
        f = np.ascontiguousarray(f, np.float32)
        lo, hi = np.quantile(f, [0.1, 0.9])
        out.append((f, (float(rng.uniform(-1, 1)), 0.0, 0.5), (1.0 / 7, 0.25, 0.5), sorted(rng.uniform(lo, hi, size=3).tolist())))
    return out


def single(dmc, f, origin, delta, iso, flags, refinement):
    dmc.extract_raw(f.ctypes.data, (f.shape[2], f.shape[1], f.shape[0]), origin, delta, iso, flags, refinement, False)
    return dmc.mesh_numpy(), dmc.last_stats.as_dict()


@pytest.mark.parametrize("flags,refinement", [(oa.PAD1 | oa.SURF_ITER | oa.ALL_POST, 0), (oa.SURF_ITER | oa.SPLIT_NM, 0),
                                              (oa.PAD1 | oa.P3X3Y_COLOR | oa.P3333 | oa.SPLIT_NM, 1)])
def test_batch_problems_equal_single_extractions(flags, refinement):
    import manifolddmc_amd as mdmc
    fields = mixed_fields(7 + refinement, 9)
    batch = mdmc.DualMarchingCubes()
    one = mdmc.DualMarchingCubes()
    try:
        print("batch of", len(fields), "grids x 3 iso, flags", flags, flush=True)
        tot = batch.extract_batch([f for f, _, _, _ in fields], [i for _, _, _, i in fields], [o for _, o, _, _ in fields],
                                  [d for _, _, d, _ in fields], flags, refinement)
        meshes = batch.batch_meshes()
        per = batch.batch_stats()
        assert len(meshes) == 3 * len(fields)
        sums = {"n_vertices": 0, "n_faces": 0, "nm_edges4": 0}
        b = 0
        for f, origin, delta, isos in fields:
            for iso in isos:
                m, st = single(one, f, origin, delta, iso, flags, refinement)
                g = meshes[b]
                errs = []
                for name in ("positions", "normals", "halfedges", "face_halfedge", "face_attr", "vertex_flags", "vertex_halfedge"):
                    e = meshcheck.exact_diff(name, getattr(g, name), getattr(m, name))
                    if e:
                        errs.append(e)
                ps = per[b].as_dict()
                for k in ("n_vertices", "n_halfedges", "n_faces", "n_dmc_vertices", "n_dmc_quads", "nm_edges4", "nm_edges3", "n_edges",
                          "quads_with_nm", "nm_vertices"):
                    if ps[k] != st[k]:
                        errs.append("stat %s: batch %s single %s" % (k, ps[k], st[k]))
                assert not errs, ("problem %d iso %g shape %s" % (b, iso, f.shape), errs)
                for k in sums:
                    sums[k] += st[k]
                b += 1
        td = tot.as_dict()
        for k in sums:
            assert td[k] == sums[k], k
    finally:
        batch.close()
        one.close()


def test_batch_with_empty_problems_and_one_grid():
    import manifolddmc_amd as mdmc
    f = np.random.default_rng(3).uniform(-1, 1, size=(9, 11, 10)).astype(np.float32)
    d = mdmc.DualMarchingCubes()
    one = mdmc.DualMarchingCubes()
    try:
        flags = oa.PAD1 | oa.SURF_ITER | oa.ALL_POST
        d.extract_batch([f, f], [[5.0, 0.0], [-5.0, 0.2]], [(0, 0, 0)] * 2, [(1, 1, 1)] * 2, flags)
        meshes = d.batch_meshes()
        assert len(meshes[0].positions) == 0 and len(meshes[2].positions) == 0
        for b, iso in ((1, 0.0), (3, 0.2)):
            m, _ = single(one, f, (0, 0, 0), (1, 1, 1), iso, flags, 0)
            assert np.array_equal(meshes[b].halfedges, m.halfedges) and np.array_equal(meshes[b].positions, m.positions)
        # nothing at all
        d.extract_batch([f], [[7.0, -7.0]], [(0, 0, 0)], [(1, 1, 1)], flags)
        assert all(len(m.positions) == 0 for m in d.batch_meshes())
    finally:
        d.close()
        one.close()


def test_batch_graph_replays_equal_eager_batches():
    """The batched form's front (padding, classification, tile scans) is captured on the second
    call with the same shapes and replayed from the third on; every call, including replays
    with new field values and iso-values (same shapes) and with other extractions in between,
    equals the same batch run with graphs off."""
    import torch

    import manifolddmc_amd as mdmc
    from manifolddmc_amd import grids
    rng = np.random.default_rng(5)
    shapes = [tuple(int(v) for v in rng.integers(12, 40, size=3)) for _ in range(6)]

    def fields(seed):
        out = []
        for i, (nz, ny, nx) in enumerate(shapes):
            f = grids.gyroid(max(nx, ny, nz), periods=2 + i % 3, noise=0.05, seed=seed + i)[:nz, :ny, :nx]
            out.append(torch.from_numpy(np.ascontiguousarray(f, np.float32)).cuda())
        return out

    flags = oa.PAD1 | oa.SURF_ITER | oa.ALL_POST
    g = mdmc.DualMarchingCubes()
    ref = mdmc.DualMarchingCubes()
    other = mdmc.DualMarchingCubes()
    ref.set_graphs(False)
    try:
        g.set_tile_shape(1)
        ref.set_tile_shape(1)
        sets = [(fields(1), [[-0.2, 0.1]] * 6), (fields(1), [[-0.2, 0.1]] * 6), (fields(1), [[0.0, 0.2]] * 6),
                (fields(9), [[-0.1, 0.3]] * 6), (fields(9), [[-0.1, 0.3]] * 6)]
        for k, (fs, isos) in enumerate(sets):
            torch.cuda.synchronize()
            gl = [(f.data_ptr(), (f.shape[2], f.shape[1], f.shape[0]), (0, 0, 0), (1, 1, 1), True) for f in fs]
            print("graph batch call", k, flush=True)
            g.extract_batch_raw(gl, isos, flags)
            a = g.mesh_numpy()
            ref.extract_batch_raw(gl, isos, flags)
            b = ref.mesh_numpy()
            for name in ("positions", "normals", "halfedges", "face_halfedge", "face_attr", "vertex_flags", "vertex_halfedge"):
                e = meshcheck.exact_diff(name, getattr(a, name), getattr(b, name))
                assert e is None, (k, e)
            # unrelated eager work between the calls
            f0 = fs[0]
            other.extract_raw(f0.data_ptr(), (f0.shape[2], f0.shape[1], f0.shape[0]), (0, 0, 0), (1, 1, 1), 0.0, flags, 0, True)
            torch.cuda.synchronize()
        assert g.graph_replays() == 4  # call 0 ran eagerly, call 1 captured the front and ran it, calls 2-4 replayed it
        assert ref.graph_replays() == 0
    finally:
        g.close()
        ref.close()
        other.close()
