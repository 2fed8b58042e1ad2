"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: a CPU restatement of the reference pipeline
used only as the checker.  Nothing in the product imports this module.
"""
import ctypes
import os
import struct

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

# option bits (same meaning as include/mdmc.h)
PAD1 = 1
SURF_ITER = 2
P3X3Y_COLOR = 4
P3X3Y_OLD = 8
P3333 = 16
SPLIT_NM = 32
DETECT_NM_VERTS = 64
ALL_POST = P3X3Y_COLOR | P3X3Y_OLD | P3333 | SPLIT_NM | DETECT_NM_VERTS

STEP_NONE, STEP_SUBDIVIDE, STEP_COLLAPSE, STEP_MERGE, STEP_TWINS, STEP_MOVE, STEP_ITERATIVE, STEP_ALL = 0, 1, 2, 3, 4, 5, 6, 100


class OrcOptions(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_int32), ("refinement", ctypes.c_int32),
                ("split_step", ctypes.c_int32), ("max_split_iters", ctypes.c_int32)]


class OrcStats(ctypes.Structure):
    _fields_ = [("n_vertices", ctypes.c_int64), ("n_halfedges", ctypes.c_int64), ("n_faces", ctypes.c_int64),
                ("n_dmc_vertices", ctypes.c_int64), ("n_dmc_quads", ctypes.c_int64),
                ("nm_edges4", ctypes.c_int64), ("nm_edges3", ctypes.c_int64), ("n_edges", ctypes.c_int64),
                ("nm_vertices", ctypes.c_int64), ("quads_with_nm", ctypes.c_int64 * 5),
                ("split_iterations", ctypes.c_int64), ("anomalies", ctypes.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_extract.restype = ctypes.c_void_p
        L.orc_extract.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                  ctypes.POINTER(OrcOptions), ctypes.c_int32, ctypes.POINTER(OrcStats)]
        L.orc_get_mesh.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 7
        L.orc_get_positions_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_get_grid.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_simplify_quads.restype = ctypes.c_void_p
        L.orc_simplify_quads.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(OrcStats)]
        L.orc_set_plant.restype = ctypes.c_int32
        L.orc_set_plant.argtypes = [ctypes.c_int32]
        L.orc_cell_polygon.restype = ctypes.c_int32
        L.orc_cell_polygon.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleMesh:
    """The oracle's output arrays (same layout as the product's mdmc_mesh)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def extract(data, origin, delta, iso, flags=SURF_ITER | PAD1 | SPLIT_NM | DETECT_NM_VERTS, refinement=0,
            split_step=STEP_ALL, max_split_iters=64, fp64=False, want_grid=False):
    """Run the oracle on a float32 grid laid out [z][y][x] (x fastest)."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    nz, ny, nx = data.shape
    dims = np.array([nx, ny, nz], dtype=np.int32)
    org = np.asarray(origin, dtype=np.float32)
    dl = np.asarray(delta, dtype=np.float32)
    opt = OrcOptions(flags, refinement, split_step, max_split_iters)
    st = OrcStats()
    L = lib()
    h = L.orc_extract(_ptr(data), _ptr(dims), _ptr(org), _ptr(dl), ctypes.c_float(iso), ctypes.byref(opt),
                      1 if fp64 else 0, ctypes.byref(st))
    try:
        V, H, F = st.n_vertices, st.n_halfedges, st.n_faces
        pos = np.zeros((V, 3), np.float32)
        nrm = np.zeros((V, 3), np.float32)
        he = np.zeros((H, 4), np.int32)
        face_he = np.zeros(F, np.int32)
        face_attr = np.zeros(F, np.uint16)
        vflags = np.zeros(V, np.uint8)
        vhe = np.zeros(V, np.int32)
        L.orc_get_mesh(h, _ptr(pos), _ptr(nrm), _ptr(he), _ptr(face_he), _ptr(face_attr), _ptr(vflags), _ptr(vhe))
        out = OracleMesh(pos=pos, nrm=nrm, he=he, face_he=face_he, face_attr=face_attr, vflags=vflags, vhe=vhe,
                         stats={k: (list(getattr(st, k)) if k == "quads_with_nm" else getattr(st, k))
                                for k, _ in OrcStats._fields_})
        if fp64:
            p64 = np.zeros((V, 3), np.float64)
            n64 = np.zeros((V, 3), np.float64)
            L.orc_get_positions_f64(h, _ptr(p64), _ptr(n64))
            out.pos64, out.nrm64 = p64, n64
        if want_grid:
            gd = np.zeros(3, np.int32)
            go = np.zeros(3, np.float32)
            gdl = np.zeros(3, np.float32)
            L.orc_get_grid(h, _ptr(gd), _ptr(go), _ptr(gdl), None)
            vals = np.zeros(int(gd[0]) * int(gd[1]) * int(gd[2]), np.float32)
            L.orc_get_grid(h, _ptr(gd), _ptr(go), _ptr(gdl), _ptr(vals))
            out.grid = vals.reshape(int(gd[2]), int(gd[1]), int(gd[0]))
            out.grid_origin, out.grid_delta = go, gdl
        return out
    finally:
        L.orc_free(h)


def _mesh_from_handle(L, h, st):
    V, H, F = st.n_vertices, st.n_halfedges, st.n_faces
    pos = np.zeros((V, 3), np.float32)
    nrm = np.zeros((V, 3), np.float32)
    he = np.zeros((H, 4), np.int32)
    face_he = np.zeros(F, np.int32)
    face_attr = np.zeros(F, np.uint16)
    vflags = np.zeros(V, np.uint8)
    vhe = np.zeros(V, np.int32)
    L.orc_get_mesh(h, _ptr(pos), _ptr(nrm), _ptr(he), _ptr(face_he), _ptr(face_attr), _ptr(vflags), _ptr(vhe))
    return OracleMesh(pos=pos, nrm=nrm, he=he, face_he=face_he, face_attr=face_attr, vflags=vflags, vhe=vhe,
                      stats={k: (list(getattr(st, k)) if k == "quads_with_nm" else getattr(st, k)) for k, _ in OrcStats._fields_})


def simplify_quads(pos, nrm, quads, attr, flags):
    """The oracle's halfedge initialize plus the simplifications selected in flags
    (P3X3Y_COLOR, P3X3Y_OLD, P3333) on a caller-built quad mesh."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
    quads = np.ascontiguousarray(quads, np.int32).reshape(-1, 4)
    attr = np.ascontiguousarray(attr, np.uint16).reshape(-1)
    assert len(pos) == len(nrm) and len(attr) == len(quads)
    assert quads.size == 0 or (quads.min() >= 0 and quads.max() < len(pos))
    L = lib()
    st = OrcStats()
    h = L.orc_simplify_quads(len(pos), _ptr(pos), _ptr(nrm), len(quads), _ptr(quads), _ptr(attr), flags, ctypes.byref(st))
    try:
        return _mesh_from_handle(L, h, st)
    finally:
        L.orc_free(h)


# planted mistakes of the oracle (orc::PLANT_* in oracle/mdmc_oracle.cpp)
PLANT_SWAP_SURVIVOR, PLANT_COLOR_LT, PLANT_NO_RENORM, PLANT_FAR_CORNER = 1, 2, 3, 4
PLANT_NO_ALL_FOUR, PLANT_NO_NEIGHBOR_NM, PLANT_OLD_NO_COUNT, PLANT_MID_XZ = 5, 6, 7, 8


class planted:
    """with planted(oa.PLANT_...): the oracle runs with that one mistake."""

    def __init__(self, which):
        self.which = which

    def __enter__(self):
        self.old = lib().orc_set_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(self.old)
        return False


def cell_polygon(values, iso, use_decider):
    s = np.ascontiguousarray(values, dtype=np.float32)
    sizes = np.zeros(4, np.int32)
    edges = np.zeros(12, np.int32)
    n = lib().orc_cell_polygon(_ptr(s), ctypes.c_float(iso), 1 if use_decider else 0, _ptr(sizes), _ptr(edges))
    out, k = [], 0
    for c in range(n):
        out.append([int(e) for e in edges[k:k + sizes[c]]])
        k += sizes[c]
    return out


# ----------------------------------------------------------------------------
# reference file formats (UniformGrid.inl:512-571 grid reader; Mesh.cu:446-506 .obj writer)
def read_bin(path):
    """Grid file: int32 dims[3], float32 deltas[3], then uint16 values x fastest.

    Returns (values[z][y][x] float32, origin float32[3], deltas float32[3]) with
    origin = -0.5 * (dims - 1) * deltas in float32, as the reference computes it."""
    b = open(path, "rb").read()
    dims = struct.unpack("<3i", b[:12])
    deltas = np.array(struct.unpack("<3f", b[12:24]), dtype=np.float32)
    n = dims[0] * dims[1] * dims[2]
    vals = np.frombuffer(b[24:24 + 2 * n], dtype="<u2").astype(np.float32)
    origin = (np.float32(-0.5) * (np.array(dims, dtype=np.int32) - 1).astype(np.float32)) * deltas
    return vals.reshape(dims[2], dims[1], dims[0]), origin.astype(np.float32), deltas


def read_obj(path):
    """Reads v / vn / f a//a ... lines; faces returned 0-based."""
    vs, ns, fs = [], [], []
    for line in open(path):
        p = line.split()
        if not p:
            continue
        if p[0] == "v":
            vs.append([float(x) for x in p[1:4]])
        elif p[0] == "vn":
            ns.append([float(x) for x in p[1:4]])
        elif p[0] == "f":
            fs.append([int(t.split("/")[0]) - 1 for t in p[1:5]])
    return np.array(vs, np.float64), np.array(ns, np.float64), np.array(fs, np.int64).reshape(-1, 4)
