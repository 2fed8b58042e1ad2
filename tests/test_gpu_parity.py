"""HIP library against the CPU oracle: identical arrays (bit-exact floats, exact ids)."""
import numpy as np
import pytest

import fixtures
import meshcheck
import oracle_api as oa

pytestmark = pytest.mark.gpu

_dmc = None


def gpu():
    global _dmc
    if _dmc is None:
        import manifolddmc_amd as mdmc
        _dmc = mdmc.DualMarchingCubes()
    return _dmc


def run_gpu(vals, origin, delta, iso, flags, refinement=0, split_step=oa.STEP_ALL):
    d = gpu()
    d.splitting_step = split_step
    v = np.ascontiguousarray(vals, np.float32)
    nz, ny, nx = v.shape
    st = d.extract_raw(v.ctypes.data, (nx, ny, nz), origin, delta, iso, flags, refinement, False)
    return d.mesh_numpy(), st.as_dict()


def assert_same(vals, origin, delta, iso, flags, refinement=0, split_step=oa.STEP_ALL, label=""):
    print("case", label, "flags", flags, "shape", np.shape(vals), flush=True)
    g, gst = run_gpu(vals, origin, delta, iso, flags, refinement, split_step)
    o = oa.extract(vals, origin, delta, iso, flags=flags, refinement=refinement, split_step=split_step)
    errs = []
    for name, a, b in [("positions", g.positions, o.pos), ("normals", g.normals, o.nrm), ("halfedges", g.halfedges, o.he),
                       ("face_halfedge", g.face_halfedge, o.face_he), ("face_attr", g.face_attr, o.face_attr),
                       ("vertex_flags", g.vertex_flags, o.vflags), ("vertex_halfedge", g.vertex_halfedge, o.vhe)]:
        e = meshcheck.exact_diff(name, a, b)
        if e:
            errs.append(e)
    for k in ("n_vertices", "n_halfedges", "n_faces", "n_dmc_vertices", "n_dmc_quads", "nm_edges4", "nm_edges3", "n_edges",
              "nm_vertices", "quads_with_nm", "split_iterations", "anomalies"):
        if gst[k] != o.stats[k]:
            errs.append("stat %s: gpu %s oracle %s" % (k, gst[k], o.stats[k]))
    # no single-pass scan or compaction tile waited past its spin limit on the tiles before it
    if gst["lookback_fallbacks"] != 0:
        errs.append("look-back fallbacks: %d" % gst["lookback_fallbacks"])
    assert not errs, "\n".join(errs)
    return g, gst


@pytest.mark.parametrize("setting,problem", fixtures.cases())
def test_fixture_cases_match_oracle(setting, problem):
    vals, origin, deltas, iso = fixtures.load_problem(problem)
    kw = fixtures.settings_for(setting, problem)
    g, _ = assert_same(vals, origin, deltas, iso, kw["flags"], split_step=kw["split_step"], label=setting + "/" + problem)
    res = fixtures.compare_with_obj(g.positions, g.faces(), fixtures.obj_path(setting, problem), nrm=g.normals)
    assert res == {"n_vertices": 0, "n_faces": 0, "bad_position": 0, "bad_normal": 0, "face_mismatch": 0, "unreferenced_delta": 0}, res


@pytest.mark.parametrize("step", [0, 1, 2, 3, 4, 5, 6])
def test_split_steps_match_oracle(step):
    vals, origin, deltas, iso = fixtures.load_problem("Iterative-Fix")
    assert_same(vals, origin, deltas, iso, oa.SURF_ITER | oa.PAD1 | oa.SPLIT_NM | oa.DETECT_NM_VERTS, split_step=step, label="step%d" % step)


FLAG_SETS = [
    oa.PAD1,
    oa.SURF_ITER | oa.PAD1,
    oa.SURF_ITER | oa.PAD1 | oa.SPLIT_NM | oa.DETECT_NM_VERTS,
    oa.SURF_ITER | oa.PAD1 | oa.ALL_POST,
    oa.SURF_ITER | oa.ALL_POST,
    oa.SURF_ITER | oa.PAD1 | oa.P3X3Y_COLOR,
    oa.SURF_ITER | oa.PAD1 | oa.P3X3Y_OLD,
    oa.SURF_ITER | oa.PAD1 | oa.P3333,
]


@pytest.mark.parametrize("flags", FLAG_SETS)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_small_grids(flags, seed):
    v = np.random.default_rng(100 + seed).uniform(-1, 1, size=(9, 10, 11)).astype(np.float32)
    g, st = assert_same(v, (0.1, -0.2, 0.3), (0.5, 0.25, 1.0), 0.0, flags, label="rand%d" % seed)
    if st["nm_edges3"] == 0:  # 3-face edges (only without padding) carry cyclic, non-involutive twins
        assert meshcheck.invariants(g.halfedges, g.face_halfedge, len(g.positions)) == []


@pytest.mark.parametrize("flags", FLAG_SETS[2:5])
def test_blob_field_64(flags):
    from manifolddmc_amd import grids
    v, o, d, iso = grids.workload("blobs64")
    assert_same(v, o, d, iso, flags, label="blobs64")


@pytest.mark.parametrize("flags", FLAG_SETS[2:4])
def test_nonmanifold_stress_128(flags):
    from manifolddmc_amd import grids
    v, o, d, iso = grids.workload("random128")
    g, st = assert_same(v, o, d, iso, flags, label="random128")
    assert st["split_iterations"] >= 1 and st["nm_edges4"] == 0


def test_gyroid_96_all_post():
    from manifolddmc_amd import grids
    v, o, d, iso = grids.workload("gyroid96")
    assert_same(v, o, d, iso, oa.SURF_ITER | oa.PAD1 | oa.ALL_POST, label="gyroid96")


def test_refinement_matches_oracle():
    v = np.random.default_rng(7).uniform(-1, 1, size=(6, 7, 5)).astype(np.float32)
    assert_same(v, (0, 0, 0), (1, 1, 1), 0.1, oa.SURF_ITER | oa.PAD1 | oa.SPLIT_NM | oa.DETECT_NM_VERTS, refinement=2, label="refine2")


def test_empty_and_degenerate():
    v = np.zeros((5, 5, 5), np.float32)
    g, st = run_gpu(v, (0, 0, 0), (1, 1, 1), 1.0, oa.PAD1 | oa.SURF_ITER)
    assert st["n_vertices"] == 0 and len(g.face_halfedge) == 0
    import manifolddmc_amd as mdmc
    with pytest.raises(mdmc.MdmcError):
        run_gpu(np.zeros((1, 1, 1), np.float32), (0, 0, 0), (1, 1, 1), 0.0, 0)


def nan_equal(name, a, b):
    """Floats equal bit for bit, except that any NaN matches any NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return "%s: shape %s vs %s" % (name, a.shape, b.shape)
    both = np.isnan(a) & np.isnan(b)
    same = (a.view(np.uint32) == b.view(np.uint32)) | both
    if not same.all():
        i = np.argwhere(~same)[0]
        return "%s: first difference at %s: %r vs %r" % (name, tuple(i), a[tuple(i)], b[tuple(i)])
    return None


@pytest.mark.parametrize("bad", ["nan", "inf"])
def test_non_finite_samples_match_oracle(bad):
    """NaN / +-Inf samples (the reference only zeroes NaN when it reads a file,
    UniformGrid.inl:553-560; a grid handed over in memory keeps them): the mesh
    arrays equal the oracle's, positions and normals up to NaN payloads."""
    rng = np.random.default_rng(77)
    v = rng.uniform(-1, 1, size=(12, 13, 14)).astype(np.float32)
    idx = rng.choice(v.size, size=40, replace=False)
    v.reshape(-1)[idx] = np.nan if bad == "nan" else np.where(rng.random(40) < 0.5, np.inf, -np.inf).astype(np.float32)
    for flags in (oa.PAD1 | oa.SURF_ITER, oa.PAD1 | oa.SURF_ITER | oa.ALL_POST):
        print("case non-finite", bad, "flags", flags, flush=True)
        g, gst = run_gpu(v, (0, 0, 0), (1, 1, 1), 0.0, flags)
        o = oa.extract(v, (0, 0, 0), (1, 1, 1), 0.0, flags=flags)
        errs = [e for e in (nan_equal("positions", g.positions, o.pos), nan_equal("normals", g.normals, o.nrm)) if e]
        for name, a, b in [("halfedges", g.halfedges, o.he), ("face_halfedge", g.face_halfedge, o.face_he),
                           ("face_attr", g.face_attr, o.face_attr), ("vertex_flags", g.vertex_flags, o.vflags)]:
            e = meshcheck.exact_diff(name, a, b)
            if e:
                errs.append(e)
        assert not errs, errs
