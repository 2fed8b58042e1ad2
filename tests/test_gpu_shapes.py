"""HIP library against the oracle on grid shapes and values the other tests do not reach: the
smallest grids, one- and two-cell-thin slabs along each axis, long thin rods (a tile then spans
many rows or planes), and fields whose samples sit exactly on the iso-value (the classification
test is value >= iso, UniformGrid.inl / DualMarchingCubes.cu:38-67, so ties decide the case).
"""
import numpy as np
import pytest

import oracle_api as oa
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2, 2), (3, 3, 3), (2, 2, 33), (33, 2, 2), (2, 33, 2), (3, 40, 5), (70, 3, 2), (4, 2, 300)]
FLAGS = [oa.SURF_ITER, oa.SURF_ITER | oa.PAD1 | oa.SPLIT_NM | oa.DETECT_NM_VERTS, oa.SURF_ITER | oa.PAD1 | oa.ALL_POST]


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("shape", SHAPES)
def test_thin_and_tiny_grids(shape, flags):
    nz, ny, nx = shape
    v = np.random.default_rng(nx * 10007 + ny * 101 + nz).uniform(-1, 1, size=(nz, ny, nx)).astype(np.float32)
    assert_same(v, (0.5, -1.0, 2.0), (0.25, 0.5, 1.0), 0.0, flags, label="shape %s" % (shape,))


@pytest.mark.parametrize("flags", FLAGS)
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_samples_on_the_iso_value(seed, flags):
    # values in {-1, 0, 1} (a third of the samples equal the iso-value) and in {0, 0.5} with
    # iso 0.5: every tie goes to the inside
    rng = np.random.default_rng(500 + seed)
    v = rng.integers(-1, 2, size=(12, 11, 10)).astype(np.float32)
    assert_same(v, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags, label="ternary %d" % seed)
    w = (0.5 * rng.integers(0, 2, size=(9, 13, 8))).astype(np.float32)
    assert_same(w, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.5, flags, label="binary %d" % seed)


@pytest.mark.parametrize("flags", FLAGS[1:])
def test_checkerboard_every_cell_ambiguous(flags):
    # alternating signs: every face of every cell is ambiguous, the asymptotic decider picks
    # every polygon, and the mesh is heavily non-manifold before the repair
    z, y, x = np.meshgrid(np.arange(14), np.arange(13), np.arange(12), indexing="ij")
    v = np.where((x + y + z) % 2 == 0, 1.0, -1.0).astype(np.float32)
    v *= (1.0 + 0.01 * ((x * 7 + y * 3 + z * 5) % 11)).astype(np.float32)  # break the exact symmetry of the deciders
    g, st = assert_same(v, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags, label="checkerboard")
    assert st["n_faces"] > 0


OPEN = [oa.DETECT_NM_VERTS, oa.SURF_ITER | oa.DETECT_NM_VERTS, oa.SURF_ITER | oa.SPLIT_NM | oa.DETECT_NM_VERTS]


@pytest.mark.parametrize("flags", OPEN)
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_open_fans_against_oracle(seed, flags):
    # no padding: the surface ends at the grid's faces, so many vertices have open fans, and
    # edges of three faces (twins that are not mutual) occur; the fan flags decided by the
    # walks from one outgoing halfedge must equal the oracle's per-halfedge fan keys
    rng = np.random.default_rng(700 + seed)
    v = rng.uniform(-1, 1, size=(14, 13, 12)).astype(np.float32)
    g, st = assert_same(v, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, flags, label="open fans %d" % seed)
    print("open fans seed %d flags %d: %d vertices with several fans, %d edges of 3 faces" % (seed, flags, st["nm_vertices"], st["nm_edges3"]))
    assert st["nm_vertices"] >= 0
