"""The reference's synthetic grids generated on the GPU (mdmc_generate_surface /
mdmc_generate_edge_case; UniformGrid.inl:25-269, :463-509).

  * values: the polynomial surfaces equal a float32 numpy evaluation of the same formula
    in the same operation order bit for bit; the trigonometric ones (iWP, pwHybrid,
    neovius) agree with a float64 evaluation to float32 rounding of the cosines;
  * topology, fixed by the mathematics: after the full pipeline (pad-by-1, all
    post-processes) the sphere, torus, two-holed torus, four-holed torus and genus-two
    surface are closed 2-manifolds with Euler characteristic 2, 0, -2, -6 and -2, and the
    sphere's vertices lie within a cell of radius 0.4;
  * the two 4x4x3 edge cases hold 950 exactly at their marked points and 850 elsewhere,
    and the GPU mesh of each equals the oracle's at iso 900 with and without padding.
"""
import numpy as np
import pytest

import meshcheck
import oracle_api as oa

pytestmark = pytest.mark.gpu

f32 = np.float32


def coords(dims):
    nx, ny, nz = dims
    d = [f32(2.0) / (f32(n) - f32(1.0)) for n in (nx, ny, nz)]
    x = f32(-1.0) + np.arange(nx, dtype=np.float32) * d[0]
    y = f32(-1.0) + np.arange(ny, dtype=np.float32) * d[1]
    z = f32(-1.0) + np.arange(nz, dtype=np.float32) * d[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z, d


def sq(a):
    return a * a


def torus_h(X, Y, Z, c, px, py):
    cc, a = sq(f32(px)), sq(f32(py))
    x, y, z = sq(X - f32(c[0])), sq(Y - f32(c[1])), sq(Z - f32(c[2]))
    return sq(x + y + z + cc - a) - f32(4) * cc * (x + y)


def torus_v(X, Y, Z, c, px, py):
    cc, a = sq(f32(px)), sq(f32(py))
    x, y, z = sq(X - f32(c[0])), sq(Y - f32(c[1])), sq(Z - f32(c[2]))
    return sq(x + y + z + cc - a) - f32(4) * cc * (x + z)


def numpy_surface(case, X, Y, Z):
    if case == 0:
        return X * X + Y * Y + Z * Z - f32(0.16)
    if case == 1:
        return torus_h(X, Y, Z, (0, 0, 0), 0.3, 0.15)
    if case == 2:
        return np.minimum(torus_h(X, Y, Z, (-0.38, 0, 0), 0.3, 0.15), torus_h(X, Y, Z, (0.2, 0.38, 0), 0.3, 0.15))
    if case == 3:
        v = np.minimum(torus_h(X, Y, Z, (-0.38, 0, 0), 0.3, 0.15), torus_h(X, Y, Z, (0.38, 0, 0), 0.3, 0.15))
        v = np.minimum(v, torus_v(X, Y, Z, (0, 0, -0.38), 0.3, 0.15))
        return np.minimum(v, torus_v(X, Y, Z, (0, 0, 0.38), 0.3, 0.15))
    if case == 4:
        x = (X + f32(1)) / f32(2)
        y = (Y + f32(1)) / f32(2)
        z = (Z + f32(1)) / f32(2)
        x, y, z = f32(4) * x - f32(2), f32(4) * y - f32(2), f32(4) * z - f32(2)
        t = f32(2) * y * (y * y - f32(3) * x * x) * (f32(1) - z * z)
        t = t + (x * x + y * y) * (x * x + y * y)
        return t - (f32(9) * z * z - f32(1)) * (f32(1) - z * z)
    if case == 8:
        q = lambda a: a * a * a * a
        s2 = sq(X) + sq(Y) + sq(Z)
        return q(X) + q(Y) + q(Z) + f32(-1) * s2 + f32(0) * s2 + f32(0.5)
    if case == 9:
        x, y, z = f32(1.5) * X, f32(1.5) * Y, f32(1.5) * Z
        return sq(x * x + y * y + z * z - f32(1)) - (sq(z - f32(1)) - f32(2) * x * x) * (sq(z + f32(1)) - f32(2) * y * y)
    al = {5: 5.01, 6: 1.01, 7: 1.0}[case]
    # the argument: float32 alpha * (p + 1), times the double pi, rounded to float32
    arg = lambda P: ((f32(al) * (P + f32(1))).astype(np.float64) * np.pi).astype(np.float32).astype(np.float64)
    cx, cy, cz = np.cos(arg(X)), np.cos(arg(Y)), np.cos(arg(Z))
    if case == 5:
        return cx * cy + cy * cz + cz * cx - cx * cy * cz
    if case == 6:
        return 4.0 * (cx * cy + cy * cz + cz * cx) - 3 * cx * cy * cz + 0.8
    return 3 * (cx + cy + cz) + 4 * cx * cy * cz


@pytest.mark.parametrize("case", list(range(10)))
def test_surface_values(case):
    import manifolddmc_amd as mdmc
    dims = (33, 29, 31)
    d = mdmc.DualMarchingCubes()
    try:
        v, origin, delta = d.generate_volume(dims, case)
    finally:
        d.close()
    X, Y, Z, dl = coords(dims)
    assert origin == (-1.0, -1.0, -1.0) and np.array_equal(np.array(delta, np.float32), np.array(dl, np.float32))
    ref = numpy_surface(case, X, Y, Z)
    if case in (5, 6, 7):
        # float32 cosines: a few ulp of the terms (|cos| <= 1, coefficients <= 4)
        assert np.abs(v.astype(np.float64) - ref).max() < 5e-6
    else:
        assert np.array_equal(v.view(np.uint32), ref.astype(np.float32).view(np.uint32)), np.abs(v - ref).max()


@pytest.mark.parametrize("case,chi", [(0, 2), (1, 0), (2, -2), (3, -6), (4, -2)])
def test_surface_topology(case, chi):
    import manifolddmc_amd as mdmc
    n = 64
    d = mdmc.DualMarchingCubes()
    try:
        v, origin, delta = d.generate_volume((n, n, n), case)
        st = d.extract_raw(v.ctypes.data, (n, n, n), origin, delta, 0.0, mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST, 0, False).as_dict()
        m = d.mesh_numpy()
    finally:
        d.close()
    he = m.halfedges
    assert meshcheck.invariants(he, m.face_halfedge, len(m.positions)) == []
    assert (he[:, 3] >= 0).all(), "open boundary"
    assert st["nm_edges4"] == 0 and st["nm_edges3"] == 0 and st["nm_vertices"] == 0 and st["anomalies"] == 0
    V = len(np.unique(he[:, 0]))
    E = len(he) // 2
    F = len(m.face_halfedge)
    assert V - E + F == chi, (V, E, F)
    if case == 0:
        r = np.linalg.norm(m.positions[np.unique(he[:, 0])].astype(np.float64), axis=1)
        assert np.abs(r - 0.4).max() < delta[0]


@pytest.mark.parametrize("case", [0, 1])
def test_edge_cases(case):
    import manifolddmc_amd as mdmc
    d = mdmc.DualMarchingCubes()
    try:
        v, origin, delta = d.generate_edge_case(case)
        high = {(0, 1, 2), (0, 2, 1)} | ({(1, 2, 1)} if case == 1 else set())
        for z in range(3):
            for y in range(4):
                for x in range(4):
                    assert v[z, y, x] == (950.0 if (z, y, x) in high else 850.0)
        assert origin == (-1.0, -1.0, -1.0)
        for flags in (mdmc.SURF_ITER | mdmc.SPLIT_NM | mdmc.DETECT_NM_VERTS, mdmc.PAD1 | mdmc.SURF_ITER | mdmc.ALL_POST):
            st = d.extract_raw(v.ctypes.data, (4, 4, 3), origin, delta, 900.0, flags, 0, False).as_dict()
            g = d.mesh_numpy()
            o = oa.extract(v, origin, delta, 900.0, flags=flags)
            assert np.array_equal(g.halfedges, o.he) and np.array_equal(g.positions.view(np.uint32), o.pos.view(np.uint32))
            assert np.array_equal(g.vertex_flags, o.vflags) and st["nm_vertices"] == o.stats["nm_vertices"]
    finally:
        d.close()
