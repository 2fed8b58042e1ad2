"""OriginalDMC's face exports over the GPU mesh (mdmc_mesh_face_colors,
mdmc_mesh_triangulate_maxmin).

Face colouring (FaceColoring.cu:5-141) on meshes straight from the grid pass, whose
24-colouring by grid edge is proper:
  * equals a sequential restatement (Python loops over the faces, class 5..23 in order, then
    the colour-4 pass), on small meshes;
  * every colour is in 0..4; two faces that share an edge differ, except where the reference's
    colour-4 test (which skips the fourth neighbour) or its colour-4 pass lets them coincide.
MaxMin triangulation (DualMarchingCubes.cu:51-91) with mean-ratio quality
(QualityMeasure.h:37-70), checked against their definitions in float64:
  * each face becomes two triangles on one of its diagonals, oriented like the face;
  * the chosen diagonal's smallest angle is the larger of the two (to float32 rounding of the
    angles);
  * quality = 2 sqrt(3) |e1 x e2| / (l0 + l1 + l2), in (0, 1], 1 for equilateral triangles.
"""
import numpy as np
import pytest

import fixtures

pytestmark = pytest.mark.gpu


def extract(name, flags):
    import manifolddmc_amd as mdmc
    from manifolddmc_amd import grids
    if name == "blobs24":
        v = grids.gaussian_blobs(24, n_blobs=5, seed=3)
        origin, delta, iso = (0.0, 0.0, 0.0), (1.0 / 23,) * 3, 0.5
    elif name == "random24":
        v = np.random.default_rng(9).uniform(-1, 1, size=(20, 22, 24)).astype(np.float32)
        origin, delta, iso = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0
    else:
        v, origin, delta, iso = grids.workload(name)
    d = mdmc.DualMarchingCubes()
    v = np.ascontiguousarray(v, np.float32)
    d.extract_raw(v.ctypes.data, (v.shape[2], v.shape[1], v.shape[0]), origin, delta, iso, flags, 0, False)
    return d


def neighbours(he, face_he):
    F = len(face_he)
    nb = np.full((F, 4), -1, np.int64)
    h = face_he.astype(np.int64)
    for i in range(4):
        t = he[h, 3]
        nb[:, i] = np.where(t >= 0, he[np.maximum(t, 0), 1], -1)
        h = he[h, 2]
    return nb


def colour_sequential(col24, nb):
    col = col24.astype(np.int64).copy()
    for c in range(5, 24):
        for f in np.nonzero(col24 == c)[0]:
            cs = [col[g] if g >= 0 else -1 for g in nb[f]]
            for k in range(5):
                test = cs if k < 4 else [cs[0], cs[1], cs[2], cs[2]]
                if all(x != k for x in test):
                    col[f] = k
                    break
    snap = col.copy()
    for f in np.nonzero(snap == 4)[0]:
        cs = [snap[g] if g >= 0 else -1 for g in nb[f]]
        for k in range(4):
            if all(x != k for x in cs):
                col[f] = k
                break
    return col


@pytest.mark.parametrize("name", ["blobs24", "random24"])
def test_face_colouring_matches_sequential(name):
    import manifolddmc_amd as mdmc
    d = extract(name, mdmc.PAD1 | mdmc.SURF_ITER)
    try:
        m = d.mesh_numpy()
        col = d.face_colors()
    finally:
        d.close()
    col24 = (m.face_attr & 0x1F).astype(np.int64)
    nb = neighbours(m.halfedges, m.face_halfedge)
    # the grid's edge colouring is proper on the DMC mesh
    for i in range(4):
        ok = nb[:, i] >= 0
        assert (col24[ok] != col24[nb[ok, i]]).all()
    assert np.array_equal(col.astype(np.int64), colour_sequential(col24, nb))
    assert col.max() <= 4


@pytest.mark.parametrize("name", ["blobs64", "random128"])
def test_face_colouring_five_colours(name):
    import manifolddmc_amd as mdmc
    d = extract(name, mdmc.PAD1 | mdmc.SURF_ITER)
    try:
        m = d.mesh_numpy()
        col = d.face_colors().astype(np.int64)
    finally:
        d.close()
    assert col.max() <= 4
    nb = neighbours(m.halfedges, m.face_halfedge)
    clash = 0
    for i in range(4):
        ok = nb[:, i] >= 0
        clash += int((col[ok] == col[nb[ok, i]]).sum())
    # clashes can only come from the colour-4 shortcuts, and are rare
    assert clash <= 0.001 * len(col), clash


def angles(p0, p1, p2):
    a = np.linalg.norm(p1 - p0, axis=-1)
    b = np.linalg.norm(p2 - p1, axis=-1)
    c = np.linalg.norm(p0 - p2, axis=-1)
    A = np.arccos(np.clip((b * b + c * c - a * a) / (2 * b * c), -1, 1))
    B = np.arccos(np.clip((a * a + c * c - b * b) / (2 * a * c), -1, 1))
    return np.minimum(np.minimum(A, B), np.pi - A - B)


def mean_ratio(p0, p1, p2):
    l = (np.sum((p1 - p0) ** 2, -1) + np.sum((p2 - p1) ** 2, -1) + np.sum((p0 - p2) ** 2, -1))
    return 2 * np.sqrt(3) * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=-1) / l


@pytest.mark.parametrize("name", ["blobs64", "random128"])
def test_maxmin_triangulation_and_quality(name):
    import manifolddmc_amd as mdmc
    d = extract(name, mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST)
    try:
        m = d.mesh_numpy()
        tri, q = d.triangulate_maxmin()
    finally:
        d.close()
    f = fixtures.faces_from(m.halfedges, m.face_halfedge)
    cut13 = np.stack([f[:, 1], f[:, 3], f[:, 0], f[:, 1], f[:, 2], f[:, 3]], 1)
    cut02 = np.stack([f[:, 0], f[:, 1], f[:, 2], f[:, 0], f[:, 2], f[:, 3]], 1)
    is13 = (tri == cut13).all(1)
    assert (is13 | (tri == cut02).all(1)).all()
    P = m.positions.astype(np.float64)
    m02 = np.minimum(angles(P[f[:, 0]], P[f[:, 1]], P[f[:, 2]]), angles(P[f[:, 0]], P[f[:, 2]], P[f[:, 3]]))
    m13 = np.minimum(angles(P[f[:, 1]], P[f[:, 3]], P[f[:, 0]]), angles(P[f[:, 1]], P[f[:, 2]], P[f[:, 3]]))
    chosen = np.where(is13, m13, m02)
    other = np.where(is13, m02, m13)
    good = np.isfinite(chosen) & np.isfinite(other)
    # the kernel's angles are acosf of a float32 cosine: an error of a few ulp in the cosine
    # moves an angle t by about ulp / sin(t), large for thin triangles
    tol = 1e-5 + 2e-6 / np.maximum(np.sin(np.minimum(chosen, other)), 1e-6)
    assert (chosen[good] >= other[good] - tol[good]).all(), (chosen - other + tol)[good].min()
    T = tri.reshape(-1, 3).astype(np.int64)
    expect = mean_ratio(P[T[:, 0]], P[T[:, 1]], P[T[:, 2]])
    ok = np.isfinite(expect)
    assert np.abs(q.reshape(-1)[ok] - expect[ok]).max() < 1e-4
    assert (q.reshape(-1)[ok] <= 1 + 1e-5).all() and (q.reshape(-1)[ok] >= 0).all()


def test_maxmin_quality_on_generated_sphere():
    # the sphere of radius 0.4 from the surface generator: every triangle non-degenerate
    import manifolddmc_amd as mdmc
    d = mdmc.DualMarchingCubes()
    try:
        v, origin, delta = d.generate_volume((24, 24, 24), mdmc.SurfaceCase.Sphere)
        d.extract_raw(v.ctypes.data, (24, 24, 24), origin, delta, 0.0, mdmc.PAD1 | mdmc.SURF_ITER, 0, False)
        tri, q = d.triangulate_maxmin()
    finally:
        d.close()
    assert len(q) > 0 and np.isfinite(q).all() and (q > 0).all() and (q <= 1.0 + 1e-6).all()
