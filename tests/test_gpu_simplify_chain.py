"""A chain of simplifications (P3X3Y colour, P3X3Y old, P3333) compacts once; while another
simplification follows, a plain-patch P3X3Y merge updates its two twin pairs directly and every
other merge has its region rebuilt (HalfedgeMesh.cu:859-1048 restated as one compaction).

mdmc_set_local_merges(0) sends every merge through the region rebuild.  Both paths must give
the same arrays, and the oracle's (which simplifies step by step, compacting each time), on
meshes with and without merges that the local update refuses (noisy fields have both).
"""
import numpy as np
import pytest

import oracle_api as oa
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

CHAINS = {
    "colour+old+3333": oa.P3X3Y_COLOR | oa.P3X3Y_OLD | oa.P3333,
    "colour+old": oa.P3X3Y_COLOR | oa.P3X3Y_OLD,
    "colour+3333": oa.P3X3Y_COLOR | oa.P3333,
    "old+3333": oa.P3X3Y_OLD | oa.P3333,
}


def field(name):
    from manifolddmc_amd import grids
    if name == "noise":
        return np.random.default_rng(21).uniform(-1, 1, size=(26, 24, 22)).astype(np.float32), 0.0
    if name == "blobs":
        return grids.gaussian_blobs(40, n_blobs=9, seed=5), 0.5
    if name == "gyroid":
        return grids.workload("gyroid96")[0], 0.0
    raise KeyError(name)


def extract(v, iso, flags, local):
    import manifolddmc_amd as mdmc
    d = mdmc.DualMarchingCubes()
    try:
        d.set_local_merges(local)
        v = np.ascontiguousarray(v, np.float32)
        st = d.extract_raw(v.ctypes.data, (v.shape[2], v.shape[1], v.shape[0]), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), iso, flags, 0, False)
        return d.mesh_numpy(), st.as_dict()
    finally:
        d.close()


@pytest.mark.parametrize("chain", sorted(CHAINS))
@pytest.mark.parametrize("name", ["noise", "blobs", "gyroid"])
def test_local_and_region_rebuild_agree(name, chain):
    import manifolddmc_amd as mdmc
    v, iso = field(name)
    flags = mdmc.PAD1 | mdmc.SURF_ITER | CHAINS[chain] | mdmc.SPLIT_NM | mdmc.DETECT_NM_VERTS
    a, sa = extract(v, iso, flags, True)
    b, sb = extract(v, iso, flags, False)
    for k in ("positions", "normals", "halfedges", "face_halfedge", "face_attr", "vertex_flags", "vertex_halfedge"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), (name, chain, k)
    assert sa["anomalies"] == 0 and sb["anomalies"] == 0


@pytest.mark.parametrize("chain", sorted(CHAINS))
def test_region_rebuild_matches_oracle(chain):
    import manifolddmc_amd as mdmc
    import test_gpu_parity
    v, iso = field("noise")
    flags = mdmc.PAD1 | mdmc.SURF_ITER | CHAINS[chain] | mdmc.SPLIT_NM | mdmc.DETECT_NM_VERTS
    d = test_gpu_parity.gpu()  # the context assert_same extracts with
    d.set_local_merges(False)
    try:
        assert_same(v, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), iso, flags, label="region rebuild " + chain)
    finally:
        d.set_local_merges(True)
