"""The oracle against the reference's 28 golden meshes (Data-and-Results/Base-Results)."""
import pytest

import fixtures
import oracle_api as oa


@pytest.mark.parametrize("setting,problem", fixtures.cases())
def test_oracle_reproduces_reference_mesh(setting, problem):
    vals, origin, deltas, iso = fixtures.load_problem(problem)
    m = oa.extract(vals, origin, deltas, iso, **fixtures.settings_for(setting, problem))
    res = fixtures.compare_with_obj(m.pos, fixtures.faces_from(m.he, m.face_he), fixtures.obj_path(setting, problem), nrm=m.nrm)
    assert res == {"n_vertices": 0, "n_faces": 0, "bad_position": 0, "bad_normal": 0, "face_mismatch": 0, "unreferenced_delta": 0}, res
    assert m.stats["anomalies"] == 0


def test_iterative_fix_needs_two_iterations():
    vals, origin, deltas, iso = fixtures.load_problem("Iterative-Fix")
    one = oa.extract(vals, origin, deltas, iso, **fixtures.settings_for("2_One-Iteration", "Iterative-Fix"))
    full = oa.extract(vals, origin, deltas, iso, **fixtures.settings_for("3_Fully-Fixed", "Iterative-Fix"))
    assert one.stats["nm_edges4"] == 1 and full.stats["nm_edges4"] == 0
    assert full.stats["split_iterations"] == 2


def test_boundary_non_manifold_vertex_needs_padding():
    # Data-and-Results/README.md: a non-manifold vertex at the grid boundary, removed by padding
    vals, origin, deltas, iso = fixtures.load_problem("Non-Manifold-Vertex")
    nopad = oa.extract(vals, origin, deltas, iso, **fixtures.settings_for("0_No-Fixes", "Non-Manifold-Vertex"))
    pad = oa.extract(vals, origin, deltas, iso, **fixtures.settings_for("1_Grid-Extension-Fix", "Non-Manifold-Vertex"))
    assert nopad.stats["nm_vertices"] == 1 and pad.stats["nm_vertices"] == 0


def test_twelve_merge_leaves_one_orphan():
    # README: 12 vertices merge into one; the reference mesh keeps one unreferenced vertex
    import numpy as np
    vals, origin, deltas, iso = fixtures.load_problem("12-Merge")
    m = oa.extract(vals, origin, deltas, iso, **fixtures.settings_for("3_Fully-Fixed", "12-Merge"))
    f = fixtures.faces_from(m.he, m.face_he)
    assert len(m.pos) - len(np.unique(f)) == 1
