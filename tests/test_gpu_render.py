"""Rendering export (HalfedgeMesh.cu:690-760) and the .obj orphan lines (Mesh.cu:446-506).

  * mdmc_mesh_triangles: two triangles (x,y,z), (z,w,x) per face, the face's vertices
    walked from its first halfedge -- derived here from the oracle's halfedge arrays;
  * mdmc_mesh_render_vertices: per vertex position, normal (the oracle's) and the colour
    HSV(hue, 0.8, 0.8) for its flags, checked with Python's colorsys; vertices no halfedge
    starts at stay all zero;
  * .obj files: unreferenced vertices print as `v 0 0 0` / `vn 0 0 0`, as in the reference's
    Data-and-Results meshes (e.g. 0_No-Fixes/1-Edge.obj vertices 33 and 35).
"""
import colorsys

import numpy as np
import pytest

import fixtures
import oracle_api as oa

pytestmark = pytest.mark.gpu

# hue by vertex flags (insertVertices, HalfedgeMesh.cu:731-751), later rules overriding earlier
NM, WEIRD, FANS = 1, 4, 8


def hue_of(flags):
    h = 180.0
    if flags & NM:
        h = 0.0
    if flags & WEIRD:
        h = 60.0
    if flags & FANS:
        h = 270.0
    if (flags & WEIRD) and (flags & NM):
        h = 300.0
    if (flags & NM) and (flags & FANS):
        h = 240.0
    return h


CASES = [(s, p) for s in ("0_No-Fixes", "3_Fully-Fixed") for p in ("1-Edge", "Non-Manifold-Vertex", "12-Merge", "Iterative-Fix")]


def extract_case(setting, problem):
    import manifolddmc_amd as mdmc
    vals, origin, deltas, iso = fixtures.load_problem(problem)
    kw = fixtures.settings_for(setting, problem)
    d = mdmc.DualMarchingCubes()
    d.splitting_step = kw["split_step"]
    v = np.ascontiguousarray(vals, np.float32)
    d.extract_raw(v.ctypes.data, (v.shape[2], v.shape[1], v.shape[0]), origin, deltas, iso, kw["flags"], 0, False)
    o = oa.extract(vals, origin, deltas, iso, **kw)
    return d, o


@pytest.mark.parametrize("setting,problem", CASES)
def test_triangles_and_render_vertices(setting, problem):
    print("render export", setting, problem, flush=True)
    d, o = extract_case(setting, problem)
    try:
        tri = d.triangles()
        rv = d.render_vertices()
    finally:
        d.close()
    faces = fixtures.faces_from(o.he, o.face_he)
    expect = np.stack([faces[:, 0], faces[:, 1], faces[:, 2], faces[:, 2], faces[:, 3], faces[:, 0]], axis=1).astype(np.uint32)
    assert np.array_equal(tri, expect)
    used = np.zeros(len(o.pos), bool)
    used[o.he[:, 0]] = True
    assert np.array_equal(rv[used, 0:3], o.pos[used]) and np.array_equal(rv[used, 3:6], o.nrm[used])
    assert (rv[~used] == 0).all()
    for v in np.nonzero(used)[0]:
        rgb = colorsys.hsv_to_rgb(hue_of(int(o.vflags[v])) / 360.0, 0.8, 0.8)
        assert np.allclose(rv[v, 6:9], rgb, atol=1e-6) and rv[v, 9] == 1.0, (v, o.vflags[v], rv[v, 6:10], rgb)


@pytest.mark.parametrize("setting,problem", [("0_No-Fixes", "1-Edge"), ("3_Fully-Fixed", "12-Merge")])
def test_obj_orphans_print_as_zero(tmp_path, setting, problem):
    d, _ = extract_case(setting, problem)
    try:
        m = d.mesh_numpy()
    finally:
        d.close()
    path = str(tmp_path / "mesh.obj")
    m.write_obj(path)
    ov, on, of = oa.read_obj(path)
    rv, rn, rf = oa.read_obj(fixtures.obj_path(setting, problem))
    orphans = np.setdiff1d(np.arange(len(ov)), np.unique(of))
    ref_orphans = np.setdiff1d(np.arange(len(rv)), np.unique(rf))
    assert len(orphans) == len(ref_orphans) > 0
    assert (ov[orphans] == 0).all() and (on[orphans] == 0).all()
    assert (rv[ref_orphans] == 0).all() and (rn[ref_orphans] == 0).all()
