"""The splitter's merge phase (mergeHalfedges, ManifoldSplitter.cu:1021-1067) in its one-thread ascending pass.

The merges run in parallel and fall back to that pass only when two of them conflict, which no
bench or fixture field triggers.  mdmc_set_sequential_merges(1) makes every split iteration take
it; the arrays must still equal the oracle's (which merges one halfedge after another in
ascending order, as the pass does) on the split fixtures, every split step and the
non-manifold stress field, and the statistics must report one sequential pass per iteration that
reached the merge step (steps 4 and up).
"""
import numpy as np
import pytest

import fixtures
import oracle_api as oa
from test_gpu_parity import assert_same, gpu

pytestmark = pytest.mark.gpu

SPLIT = oa.SURF_ITER | oa.PAD1 | oa.SPLIT_NM | oa.DETECT_NM_VERTS


def sequential(fn):
    d = gpu()  # the context assert_same extracts with
    d.set_sequential_merges(True)
    try:
        return fn()
    finally:
        d.set_sequential_merges(False)


@pytest.mark.parametrize("step", [3, 4, 6])
def test_split_steps_sequential_merges(step):
    vals, origin, deltas, iso = fixtures.load_problem("Iterative-Fix")
    g, st = sequential(lambda: assert_same(vals, origin, deltas, iso, SPLIT, split_step=step, label="sequential step%d" % step))
    expect = st["split_iterations"] if step >= 4 else 0
    assert st["split_sequential_merges"] == expect, st


@pytest.mark.parametrize("setting,problem", [c for c in fixtures.cases() if c[0].startswith(("2_", "3_"))])
def test_split_fixtures_sequential_merges(setting, problem):
    vals, origin, deltas, iso = fixtures.load_problem(problem)
    kw = fixtures.settings_for(setting, problem)
    g, st = sequential(lambda: assert_same(vals, origin, deltas, iso, kw["flags"], split_step=kw["split_step"], label="sequential " + problem))
    res = fixtures.compare_with_obj(g.positions, g.faces(), fixtures.obj_path(setting, problem), nrm=g.normals)
    assert res == {"n_vertices": 0, "n_faces": 0, "bad_position": 0, "bad_normal": 0, "face_mismatch": 0, "unreferenced_delta": 0}, res


def test_stress_field_sequential_merges():
    from manifolddmc_amd import grids
    v, o, d, iso = grids.workload("random128")
    g, st = sequential(lambda: assert_same(v, o, d, iso, SPLIT, label="sequential random128"))
    assert st["split_iterations"] >= 1 and st["split_sequential_merges"] == st["split_iterations"], st


def test_parallel_merges_report_their_fallbacks():
    """Without the switch the count is the number of iterations whose parallel merge found a
    conflict (printed for the record), at most one per iteration."""
    from manifolddmc_amd import grids
    v, o, d, iso = grids.workload("random128")
    g, st = assert_same(v, o, d, iso, SPLIT, label="parallel random128")
    print("random128: split iterations", st["split_iterations"], "with a sequential merge pass", st["split_sequential_merges"])
    assert 0 <= st["split_sequential_merges"] <= st["split_iterations"], st
