"""The 512^3 gyroid workload (8 periods, +0.1 N(0,1), iso 0, pad-by-1, all
post-processes) at full size, where the oracle is too slow to compare against:
size-independent properties of the result.

  * halfedge structure: twin is an involution between opposite halfedges, next
    cycles close after 4 steps, faces point at their own halfedges;
  * the split leaves no edge shared by 4 faces and the statistics match the arrays;
  * two extractions give identical arrays, and so do two different grid-pass tile
    shapes (the tile decomposition must not show in the result).
"""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 512


@pytest.fixture(scope="module")
def field():
    import torch
    import bench
    f = bench.make_field(N)
    return torch.from_numpy(f).cuda()


def extract(t, shape):
    import manifolddmc_amd as mdmc
    d = mdmc.DualMarchingCubes()
    try:
        d.set_tile_shape(shape)
        st = d.extract_raw(t.data_ptr(), (N, N, N), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST, 0, True)
        return d.mesh_numpy(), st.as_dict()
    finally:
        d.close()


def digest(m):
    h = hashlib.sha256()
    for name in ("positions", "normals", "halfedges", "face_halfedge", "face_attr", "vertex_flags", "vertex_halfedge"):
        h.update(np.ascontiguousarray(getattr(m, name)).tobytes())
    return h.hexdigest()


def test_fullsize_structure_and_statistics(field):
    print("gyroid 512^3 all post-processes, shape 1", flush=True)
    m, st = extract(field, 1)
    he = np.asarray(m.halfedges)
    H = len(he)
    V = len(m.positions)
    F = len(m.face_halfedge)
    assert H == st["n_halfedges"] and V == st["n_vertices"] and F == st["n_faces"] and H == 4 * F
    origin, face, nxt, twin = he[:, 0], he[:, 1], he[:, 2], he[:, 3]
    assert origin.min() >= 0 and origin.max() < V and face.min() >= 0 and face.max() < F
    assert nxt.min() >= 0 and nxt.max() < H and twin.min() >= -1 and twin.max() < H
    idx = np.arange(H)
    # next cycles of length 4 within one face
    n4 = nxt[nxt[nxt[nxt]]]
    assert np.array_equal(n4, idx)
    assert np.array_equal(face[nxt], face)
    fh = np.asarray(m.face_halfedge)
    assert np.array_equal(face[fh], np.arange(F))
    # twins: an involution between halfedges running the opposite way
    has = twin >= 0
    t = twin[has]
    assert np.array_equal(twin[t], idx[has])
    assert np.array_equal(origin[t], origin[nxt[has]])
    assert np.array_equal(origin[nxt[t]], origin[has])
    # the iterative split leaves no edge with 4 faces, and nothing the reference leaves
    # undefined happened (out-of-range ids, fan walks past their guard, oversized buckets)
    assert st["nm_edges4"] == 0
    assert st["anomalies"] == 0
    assert st["lookback_fallbacks"] == 0  # no scan tile waited past its spin limit
    assert st["n_edges"] > 0 and st["n_dmc_quads"] > F


def test_fullsize_deterministic_and_independent_of_tile_shape(field):
    print("gyroid 512^3: shapes 1, 1, 5, 7, 8", flush=True)
    a, _ = extract(field, 1)
    b, _ = extract(field, 1)
    c, _ = extract(field, 5)
    d, _ = extract(field, 7)
    e, _ = extract(field, 8)
    da, db, dc, dd, de = digest(a), digest(b), digest(c), digest(d), digest(e)
    assert da == db
    assert da == dc
    assert da == dd and da == de
