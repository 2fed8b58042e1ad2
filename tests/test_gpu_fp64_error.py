"""Error of the GPU's FP32 positions and normals against the FP64 oracle (the same
algorithm evaluated in double precision), on the Gaussian-blob 64^3 field, the
non-manifold stress field (32^3 uniform lattice upsampled x4 to 128^3) and a 32-plane slab
of the 512^3 gyroid (planes 0..31, 8 periods, +0.1 N(0,1)).

The FP32 and FP64 runs must give the same topology (halfedge arrays equal).  The error is
stated per stage, because the stages differ in how far a rounding difference can travel:
  * dual vertices (no post-process): relative error |p32 - p64|inf / |p64|inf of every
    vertex at most 1e-5 (the bar for dual-vertex positions);
  * after the simplifications: midpoints of FP32 vertices, the same order of error;
  * after the non-manifold split: a vertex the splitter moves to the surface inside its
    cell follows the move-to-surface iteration of CellIntersection.inl:12-122, whose
    branches (cell-wall hits, bracketing tests) can resolve differently in FP64; such a
    vertex can land elsewhere on the surface in its cell.  Bounded here by half a cell.
The measured maxima are printed for the record.
"""
import numpy as np
import pytest

import oracle_api as oa

pytestmark = pytest.mark.gpu

ALL = oa.PAD1 | oa.SURF_ITER | oa.ALL_POST
DMC = oa.PAD1 | oa.SURF_ITER
SIMPLIFY = oa.PAD1 | oa.SURF_ITER | oa.P3X3Y_COLOR | oa.P3X3Y_OLD | oa.P3333


def workload(name):
    from manifolddmc_amd import grids
    if name == "gyroid512_slab32":
        return grids.gyroid(512, z0=0, z1=32), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0
    return grids.workload(name)


def gpu(v, origin, delta, iso, flags):
    import manifolddmc_amd as mdmc
    d = mdmc.DualMarchingCubes()
    try:
        v = np.ascontiguousarray(v, np.float32)
        d.extract_raw(v.ctypes.data, (v.shape[2], v.shape[1], v.shape[0]), origin, delta, iso, flags, 0, False)
        return d.mesh_numpy()
    finally:
        d.close()


def errors(g, o64, delta):
    e = np.abs(g.positions.astype(np.float64) - o64.pos64)
    rel = e.max(axis=1) / np.maximum(np.abs(o64.pos64).max(axis=1), np.float64(np.min(delta)))
    cells = (e / np.asarray(delta, np.float64)).max()
    nrm = np.abs(g.normals.astype(np.float64) - o64.nrm64).max()
    return float(rel.max()), float(cells), float(nrm)


@pytest.mark.parametrize("name", ["blobs64", "random128", "gyroid512_slab32"])
def test_fp32_positions_against_fp64_oracle(name):
    v, origin, delta, iso = workload(name)
    rows = []
    for label, flags, bound_rel, bound_cells in (("dual vertices", DMC, 1e-5, 1e-3), ("simplified", SIMPLIFY, 1e-5, 1e-3),
                                                 ("all post-processes", ALL, 1e-2, 0.5)):
        print("fp64 error:", name, label, flush=True)
        g = gpu(v, origin, delta, iso, flags)
        o = oa.extract(v, origin, delta, iso, flags=flags, fp64=True)
        assert np.array_equal(g.halfedges, o.he), (name, label, "topology differs from FP64")
        rel, cells, nrm = errors(g, o, delta)
        rows.append((label, len(g.positions), rel, cells, nrm))
        assert rel <= bound_rel and cells <= bound_cells, (name, label, rel, cells)
    for r in rows:
        print("%s %-20s V=%-8d max rel %.3g  max %.3g cells  normal max abs %.3g" % ((name,) + r))
