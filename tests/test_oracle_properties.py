"""Checks that pin the oracle where the reference's fixtures do not: the polygon table
against the asymptotic-decider algorithm, closed-form fields, halfedge invariants,
the FP64 variant, padding and refinement."""
import numpy as np
import pytest

import meshcheck
import oracle_api as oa


def _cyc(c):
    k = c.index(min(c))
    return tuple(c[k:] + c[:k])


def test_table_polygons_match_decider_algorithm():
    # every corner-sign case: the table contours are the cycles the asymptotic algorithm
    # builds, with the same orientation (start and order may differ)
    for case in range(1, 255):
        s = np.array([1.0 if (case >> k) & 1 else -1.0 for k in range(8)], np.float32)
        table = oa.cell_polygon(s, 0.0, False)
        algo = oa.cell_polygon(s, 0.0, True)
        assert sorted(map(_cyc, table)) == sorted(map(_cyc, algo)), case


def test_single_corner_case_contour():
    # corner 0 inside: one triangle through edges 0 (x), 3 (y), 8 (z)
    s = np.array([1, -1, -1, -1, -1, -1, -1, -1], np.float32)
    assert [sorted(c) for c in oa.cell_polygon(s, 0.0, False)] == [[0, 3, 8]]


def _plane(n, a):
    x = np.arange(n, dtype=np.float32)
    return np.broadcast_to(x[None, None, :] - np.float32(a), (n, n, n)).copy()


@pytest.mark.parametrize("surf_iter", [False, True])
def test_plane_gives_regular_quad_grid(surf_iter):
    n, a = 9, 3.3
    vals = _plane(n, a)
    flags = oa.SURF_ITER if surf_iter else 0
    m = oa.extract(vals, (0, 0, 0), (1, 1, 1), 0.0, flags=flags)
    assert len(m.pos) == (n - 1) ** 2 and len(m.face_he) == (n - 2) ** 2
    assert np.allclose(m.pos[:, 0], a, atol=1e-5)
    frac = np.sort(np.unique(np.round(m.pos[:, 1] - np.floor(m.pos[:, 1]), 6)))
    assert np.allclose(frac, [0.5])
    assert np.allclose(m.nrm, [1, 0, 0], atol=1e-6)
    assert meshcheck.invariants(m.he, m.face_he, len(m.pos)) == []
    # orientation: the face normal (from the quad's vertex order) points along +x or -x consistently
    f = np.asarray([[m.he[h, 0], m.he[m.he[h, 2], 0], m.he[m.he[m.he[h, 2], 2], 0]] for h in m.face_he])
    nrm = np.cross(m.pos[f[:, 1]] - m.pos[f[:, 0]], m.pos[f[:, 2]] - m.pos[f[:, 0]])
    assert (np.sign(nrm[:, 0]) == np.sign(nrm[0, 0])).all()


def test_sphere_vertices_near_surface():
    n = 24
    t = np.arange(n, dtype=np.float64) / (n - 1)
    r = 0.31
    d = np.sqrt((t[:, None, None] - 0.5) ** 2 + (t[None, :, None] - 0.5) ** 2 + (t[None, None, :] - 0.5) ** 2) - r
    m = oa.extract(d.astype(np.float32), (0, 0, 0), (1.0 / (n - 1),) * 3, 0.0, flags=oa.SURF_ITER | oa.PAD1 | oa.SPLIT_NM | oa.DETECT_NM_VERTS)
    rad = np.linalg.norm(m.pos - 0.5, axis=1)
    assert np.abs(rad - r).max() < 0.5 / (n - 1)
    radial = (m.pos - 0.5) / rad[:, None]
    assert (np.sum(radial * m.nrm, axis=1) > 0.95).all()
    # closed surface: every edge has two faces, Euler characteristic 2
    assert m.stats["nm_edges4"] == 0 and m.stats["nm_edges3"] == 0
    assert (m.he[:, 3] >= 0).all()
    V, E, F = len(m.pos), m.stats["n_edges"], len(m.face_he)
    assert V - E + F == 2
    assert meshcheck.invariants(m.he, m.face_he, V) == []


def _random_grid(n, seed):
    return np.random.default_rng(seed).uniform(-1, 1, size=(n, n, n)).astype(np.float32)


@pytest.mark.parametrize("seed", range(6))
def test_invariants_after_full_pipeline(seed):
    v = _random_grid(7, seed)
    m = oa.extract(v, (0, 0, 0), (1, 1, 1), 0.0, flags=oa.SURF_ITER | oa.PAD1 | oa.ALL_POST)
    assert meshcheck.invariants(m.he, m.face_he, len(m.pos)) == []
    assert m.stats["nm_edges4"] == 0  # the iterative split removes every 4-face edge
    assert m.stats["anomalies"] == 0


def _closed_blob(n=20, seed=0):
    from manifolddmc_amd import grids
    return grids.gaussian_blobs(n, n_blobs=3, seed=seed)


@pytest.mark.parametrize("flag", [oa.P3X3Y_COLOR, oa.P3X3Y_OLD, oa.P3333])
def test_simplification_keeps_closed_surface_topology(flag):
    v = _closed_blob()
    base = oa.extract(v, (0, 0, 0), (1, 1, 1), 0.5, flags=oa.SURF_ITER | oa.PAD1)
    simp = oa.extract(v, (0, 0, 0), (1, 1, 1), 0.5, flags=oa.SURF_ITER | oa.PAD1 | flag)
    chi = lambda m: len(np.unique(m.he[:, 0])) - m.stats["n_edges"] + len(m.face_he)
    assert chi(simp) == chi(base)
    assert len(simp.face_he) < len(base.face_he)  # something was simplified
    assert meshcheck.invariants(simp.he, simp.face_he, len(simp.pos)) == []
    assert (simp.he[:, 3] >= 0).all()


def test_fp64_close_to_fp32():
    v = _closed_blob()
    a = oa.extract(v, (0, 0, 0), (1, 1, 1), 0.5, flags=oa.PAD1)
    b = oa.extract(v, (0, 0, 0), (1, 1, 1), 0.5, flags=oa.PAD1, fp64=True)
    assert np.array_equal(a.he, b.he)
    assert np.abs(a.pos.astype(np.float64) - b.pos64).max() < 1e-5


def test_padding_duplicates_border_and_refinement_interpolates():
    v = _random_grid(5, 11)
    m = oa.extract(v, (0.5, -1, 2), (0.25, 0.5, 1.0), 0.0, flags=oa.PAD1, want_grid=True)
    g = m.grid
    assert g.shape == (7, 7, 7)
    assert np.array_equal(g[1:-1, 1:-1, 1:-1], v)
    assert np.array_equal(g[0, 1:-1, 1:-1], v[0]) and np.array_equal(g[1:-1, 1:-1, -1], v[:, :, -1])
    assert np.allclose(m.grid_origin, [0.25, -1.5, 1.0]) and np.allclose(m.grid_delta, [0.25, 0.5, 1.0])
    # refinement 1 of a trilinear (here linear) field reproduces it at the new points
    x = np.arange(4, dtype=np.float32)
    lin = (x[None, None, :] + 2 * x[None, :, None] - 3 * x[:, None, None]).astype(np.float32)
    r = oa.extract(lin, (0, 0, 0), (1, 1, 1), 0.25, flags=0, refinement=1, want_grid=True)
    xr = np.arange(7, dtype=np.float32) / 2
    expect = xr[None, None, :] + 2 * xr[None, :, None] - 3 * xr[:, None, None]
    assert r.grid.shape == (7, 7, 7) and np.allclose(r.grid, expect, atol=1e-5)
    assert np.allclose(r.grid_delta, [0.5, 0.5, 0.5])


def test_empty_surface():
    v = np.zeros((6, 6, 6), np.float32)
    m = oa.extract(v, (0, 0, 0), (1, 1, 1), 1.0)
    assert len(m.pos) == 0 and len(m.face_he) == 0


def _zero_weight_blend(s):
    """trilinear(s, (0, 0, 0)) in float32, in the operation order of the refinement blend
    (Interpolation.inl:48-56): lerp in x, then y, then z, each a + w * (b - a) / (1 - w) * a + w * b."""
    f = np.float32
    z, one = f(0), f(1)
    lx = [s[i] + z * (s[i + 1] - s[i]) for i in (0, 2, 4, 6)]
    ly = [(one - z) * lx[0] + z * lx[1], (one - z) * lx[2] + z * lx[3]]
    return (one - z) * ly[0] + z * ly[1]


def test_zero_weight_blend_of_plain_values_is_the_first_corner():
    # the padding pass without refinement copies the first corner unless some input value is
    # not plain (|v| >= 2^126 or -0); this is the arithmetic fact it relies on
    rng = np.random.default_rng(4)
    scales = np.float32([1e-42, 1e-30, 1.0, 1e20, 8.4e37])  # a subnormal up to just under 2^126
    with np.errstate(over="ignore", invalid="ignore"):
        for _ in range(4000):
            s = (rng.uniform(-1, 1, 8) * rng.choice(scales, 8)).astype(np.float32)
            s[rng.random(8) < 0.1] = np.float32(0.0)
            assert np.abs(s).max() < np.float32(2.0 ** 126)
            assert _zero_weight_blend(s).view(np.uint32) == s[0].view(np.uint32)
        # and what is not plain can change the result: -0 first corner, overflowing differences
        s = np.float32([-0.0, 1, 1, 1, 1, 1, 1, 1])
        assert _zero_weight_blend(s).view(np.uint32) == 0  # +0, not -0
        s = np.float32([3e38, -3e38, 0, 0, 0, 0, 0, 0])
        assert np.isnan(_zero_weight_blend(s))
