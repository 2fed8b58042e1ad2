"""manifolddmc_amd -- Manifold Dual Marching Cubes on AMD MI355X (gfx950).

Host-side mirror of the reference's DMC plugin interface
(MarchingCubes/VisualDMC/src/DMC/DualMarchingCubes.h:68-186) over the C ABI in
include/mdmc.h.  All computation happens in the HIP library libmdmc.so; this
module only marshals arguments.  It fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

from . import grids  # noqa: F401  (re-exported helpers)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdmc.so")

# option bits, include/mdmc.h
PAD1 = 0x01
SURF_ITER = 0x02
P3X3Y_COLOR = 0x04
P3X3Y_OLD = 0x08
P3333 = 0x10
SPLIT_NM = 0x20
DETECT_NM_VERTS = 0x40
RECORD_ADJOINT = 0x80  # keep the DMC topology for backward() (no reference counterpart)
DEFAULT_FLAGS = PAD1 | SURF_ITER | SPLIT_NM | DETECT_NM_VERTS
ALL_POST = P3X3Y_COLOR | P3X3Y_OLD | P3333 | SPLIT_NM | DETECT_NM_VERTS


class SplittingStep:
    """ManifoldSplitter.h:155-166."""
    NONE = 0
    SUBDIVIDE_QUADS = 1
    COLLAPSE_NON_MANIFOLD_EDGES = 2
    MERGE_VERTICES = 3
    UPDATE_HALFEDGE_TWINS = 4
    MOVE_VERTICES_TO_SURFACE = 5
    ENABLE_ITERATIVE_REFINEMENT = 6
    ALL = 100


# grid-pass tile shapes, include/mdmc.h (TILE_AUTO: autotuned per workload class)
TILE_AUTO = 0
TILE_SHAPES = {1: "1024 cells, XCD-chunked", 2: "2048 cells, XCD-chunked", 3: "4096 cells, XCD-chunked",
               4: "1024 cells, launch order", 5: "2048 cells, launch order", 6: "4096 cells, launch order",
               7: "row slabs of 4 rows, XCD-chunked", 8: "row slabs of 2 rows, XCD-chunked", 9: "2048 cells, work queue"}

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "CAPACITY", 4: "HASH_FULL", 5: "HIP", 6: "NO_RESULT", 7: "INTERNAL"}


class SurfaceCase:
    """UniformGrid.h:26-37: the implicit surfaces of UniformGrid.inl:25-127 (iso-value 0)."""
    Sphere, Torus, TwoHoledTorus, FourHoledTorus, GenusTwo, iWP, pwHybrid, neovius, Goursat, SteinerRoman = range(10)


class EdgeCase:
    """UniformGrid.h:39-42: 4x4x3 grids with a non-manifold vertex at iso-value 900."""
    NonManifoldVertex0, NonManifoldVertex1 = 0, 1


class MdmcError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("mdmc %s: %s" % (STATUS.get(status, status), message))
        self.status = status


class _Options(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_int32), ("refinement", ctypes.c_int32),
                ("split_step", ctypes.c_int32), ("max_split_iters", ctypes.c_int32)]


class _Grid(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 3), ("origin", ctypes.c_float * 3), ("delta", ctypes.c_float * 3),
                ("data", ctypes.c_void_p), ("data_on_device", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("n_vertices", ctypes.c_int64), ("n_halfedges", ctypes.c_int64), ("n_faces", ctypes.c_int64),
                ("n_dmc_vertices", ctypes.c_int64), ("n_dmc_quads", ctypes.c_int64),
                ("nm_edges4", ctypes.c_int64), ("nm_edges3", ctypes.c_int64), ("n_edges", ctypes.c_int64),
                ("nm_vertices", ctypes.c_int64), ("quads_with_nm", ctypes.c_int64 * 5),
                ("split_iterations", ctypes.c_int64), ("anomalies", ctypes.c_int64), ("cells", ctypes.c_int64),
                ("ms_total", ctypes.c_double), ("ms_grid", ctypes.c_double), ("ms_dmc", ctypes.c_double),
                ("ms_halfedge", ctypes.c_double), ("ms_simplify", ctypes.c_double), ("ms_fix_twins", ctypes.c_double),
                ("ms_split", ctypes.c_double), ("ms_detect", ctypes.c_double),
                ("ms_grid_pass_kernel", ctypes.c_double), ("split_sequential_merges", ctypes.c_int64),
                ("lookback_fallbacks", ctypes.c_int64)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if k == "quads_with_nm" else v
        return out


class _Info(ctypes.Structure):
    _fields_ = [("n_vertices", ctypes.c_int64), ("n_halfedges", ctypes.c_int64), ("n_faces", ctypes.c_int64)]


_lib = None

EXPORTS = ["mdmc_create", "mdmc_destroy", "mdmc_last_error", "mdmc_set_stream", "mdmc_extract",
           "mdmc_mesh_get_info", "mdmc_mesh_copy", "mdmc_mesh_triangles", "mdmc_mesh_render_vertices", "mdmc_backward",
           "mdmc_extract_batch", "mdmc_batch_offsets", "mdmc_batch_stats", "mdmc_set_tile_shape", "mdmc_get_tile_shape",
           "mdmc_generate_surface", "mdmc_generate_edge_case", "mdmc_mesh_face_colors", "mdmc_mesh_triangulate_maxmin",
           "mdmc_backward_noalloc", "mdmc_set_graphs", "mdmc_get_graph_replays", "mdmc_processed_grid",
           "mdmc_set_local_merges", "mdmc_set_sequential_merges"]


def lib():
    """The loaded libmdmc.so (built by manifolddmc_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmdmc.so is missing: build it with `python -m manifolddmc_amd.build` "
                               "(there is no CPU fallback)")
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7) but loads it
        # by another file name; loaded first, it satisfies this library's dependency and the
        # process keeps one runtime.  Loaded after ours, it would bring a second runtime that
        # finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.mdmc_create.argtypes = [ctypes.POINTER(vp), i32]
        L.mdmc_create.restype = i32
        L.mdmc_destroy.argtypes = [vp]
        L.mdmc_destroy.restype = None
        L.mdmc_last_error.argtypes = [vp]
        L.mdmc_last_error.restype = ctypes.c_char_p
        L.mdmc_set_stream.argtypes = [vp, vp]
        L.mdmc_set_stream.restype = i32
        L.mdmc_extract.argtypes = [vp, ctypes.POINTER(_Grid), ctypes.c_float, ctypes.POINTER(_Options), ctypes.POINTER(Stats)]
        L.mdmc_extract.restype = i32
        L.mdmc_mesh_get_info.argtypes = [vp, ctypes.POINTER(_Info)]
        L.mdmc_mesh_get_info.restype = i32
        L.mdmc_mesh_copy.argtypes = [vp] + [vp] * 7 + [i32]
        L.mdmc_mesh_copy.restype = i32
        L.mdmc_mesh_triangles.argtypes = [vp, vp, i32]
        L.mdmc_mesh_triangles.restype = i32
        L.mdmc_mesh_render_vertices.argtypes = [vp, vp, i32]
        L.mdmc_mesh_face_colors.argtypes = [vp, vp, i32]
        L.mdmc_mesh_face_colors.restype = i32
        L.mdmc_mesh_triangulate_maxmin.argtypes = [vp, vp, vp, i32]
        L.mdmc_mesh_triangulate_maxmin.restype = i32
        L.mdmc_generate_surface.argtypes = [vp, i32, vp, vp, i32, vp, vp]
        L.mdmc_generate_surface.restype = i32
        L.mdmc_generate_edge_case.argtypes = [vp, i32, vp, i32, vp, vp, vp]
        L.mdmc_generate_edge_case.restype = i32
        L.mdmc_processed_grid.argtypes = [vp, i32, vp, i32, vp, vp, vp]
        L.mdmc_processed_grid.restype = i32
        L.mdmc_mesh_render_vertices.restype = i32
        L.mdmc_backward.argtypes = [vp, vp, vp, i32]
        L.mdmc_backward.restype = i32
        L.mdmc_backward_noalloc.argtypes = [vp, vp, vp, i32]
        L.mdmc_set_graphs.argtypes = [vp, i32]
        L.mdmc_set_graphs.restype = i32
        L.mdmc_set_local_merges.argtypes = [vp, i32]
        L.mdmc_set_local_merges.restype = i32
        L.mdmc_set_sequential_merges.argtypes = [vp, i32]
        L.mdmc_set_sequential_merges.restype = i32
        L.mdmc_get_graph_replays.argtypes = [vp, ctypes.POINTER(i64)]
        L.mdmc_get_graph_replays.restype = i32
        L.mdmc_backward_noalloc.restype = i32
        L.mdmc_extract_batch.argtypes = [vp, ctypes.POINTER(_Grid), i32, vp, i32, ctypes.POINTER(_Options), ctypes.POINTER(Stats)]
        L.mdmc_extract_batch.restype = i32
        L.mdmc_batch_offsets.argtypes = [vp, vp, vp, vp]
        L.mdmc_batch_offsets.restype = i32
        L.mdmc_batch_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.mdmc_batch_stats.restype = i32
        L.mdmc_set_tile_shape.argtypes = [vp, i32]
        L.mdmc_set_tile_shape.restype = i32
        L.mdmc_get_tile_shape.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]
        L.mdmc_get_tile_shape.restype = i32
        _lib = L
    return _lib


class HalfedgeMesh:
    """The mesh returned by an extraction (Halfedge.h:120-210, HalfedgeVectors.h:225-282).

    positions/normals float32[V,3]; halfedges int32[H,4] = (origin, face, next, twin);
    face_halfedge int32[F]; face_attr uint16[F]; vertex_flags uint8[V]; vertex_halfedge int32[V]."""

    def __init__(self, positions, normals, halfedges, face_halfedge, face_attr, vertex_flags, vertex_halfedge):
        self.positions = positions
        self.normals = normals
        self.halfedges = halfedges
        self.face_halfedge = face_halfedge
        self.face_attr = face_attr
        self.vertex_flags = vertex_flags
        self.vertex_halfedge = vertex_halfedge

    def faces(self):
        """Vertex ids of each face, walked from its first halfedge (the order Mesh.cu:446-506 writes)."""
        he = np.asarray(self.halfedges)
        h = np.asarray(self.face_halfedge).astype(np.int64)
        out = np.empty((len(h), 4), np.int32)
        for i in range(4):
            out[:, i] = he[h, 0]
            h = he[h, 2]
        return out

    def render_positions_normals(self):
        """Positions and normals as the render vertex buffer holds them (HalfedgeMesh.cu:716-760):
        zeros for a vertex that no halfedge starts at."""
        pos = np.asarray(self.positions).copy()
        nrm = np.asarray(self.normals).copy()
        used = np.zeros(len(pos), bool)
        he = np.asarray(self.halfedges)
        if len(he):
            used[he[:, 0]] = True
        pos[~used] = 0
        nrm[~used] = 0
        return pos, nrm

    def write_obj(self, path):
        """The reference's .obj layout (Mesh.cu:446-506), written from the render vertex buffer:
        v, vn (zeros for unreferenced vertices), then f a//a b//b c//c d//d (1-based)."""
        pos, nrm = self.render_positions_normals()
        with open(path, "w") as f:
            for p in pos:
                f.write("v %g %g %g\n" % tuple(p))
            for n in nrm:
                f.write("vn %g %g %g\n" % tuple(n))
            for q in self.faces() + 1:
                f.write("f %d//%d %d//%d %d//%d %d//%d\n" % (q[0], q[0], q[1], q[1], q[2], q[2], q[3], q[3]))


def _member(name):
    """Property under one of the reference's member names that reads and writes
    the snake_case attribute `name`."""
    return property(lambda self: getattr(self, name), lambda self, v: setattr(self, name, v))


class DualMarchingCubes:
    """MI355X counterpart of DMC::DualMarchingCubes (DualMarchingCubes.h:68-186).

    The options have the reference's defaults (DualMarchingCubes.cu:110-133) and
    answer to both a snake_case name and the reference's member name
    (m_simplifyP3x3yColor, ...); the grid flags of UniformGridHost::readFromFile
    (refinement, add_zero_boundary) are passed with the grid."""

    # the reference's member names, DualMarchingCubes.h:168-185
    m_enableSurfaceToVertexIteration = _member("enable_surface_to_vertex_iteration")
    m_simplifyP3x3yColor = _member("simplify_p3x3y_color")
    m_simplifyP3x3yOld = _member("simplify_p3x3y_old")
    m_simplifyP3333 = _member("simplify_p3333")
    m_splitNonManifoldHalfedges = _member("split_non_manifold_halfedges")
    m_detectNonManifoldVertices = _member("detect_non_manifold_vertices")
    m_programmaticSplittingStep = _member("splitting_step")

    def __init__(self, device=-1):
        self.enable_surface_to_vertex_iteration = True
        self.simplify_p3x3y_color = False
        self.simplify_p3x3y_old = False
        self.simplify_p3333 = False
        self.split_non_manifold_halfedges = True
        self.detect_non_manifold_vertices = True
        self.splitting_step = SplittingStep.ALL
        self.max_split_iterations = 64
        self._ctx = ctypes.c_void_p()
        st = lib().mdmc_create(ctypes.byref(self._ctx), device)
        if st != 0:
            raise MdmcError(st, "mdmc_create failed")
        self.last_stats = None
        self.generation = 0  # counts extractions; backward() belongs to the latest one

    def close(self):
        if self._ctx:
            lib().mdmc_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def flags(self, add_zero_boundary=True):
        f = PAD1 if add_zero_boundary else 0
        if self.enable_surface_to_vertex_iteration:
            f |= SURF_ITER
        if self.simplify_p3x3y_color:
            f |= P3X3Y_COLOR
        if self.simplify_p3x3y_old:
            f |= P3X3Y_OLD
        if self.simplify_p3333:
            f |= P3333
        if self.split_non_manifold_halfedges:
            f |= SPLIT_NM
        if self.detect_non_manifold_vertices:
            f |= DETECT_NM_VERTS
        return f

    def set_tile_shape(self, shape):
        """mdmc_set_tile_shape: TILE_AUTO or one of TILE_SHAPES."""
        self._check(lib().mdmc_set_tile_shape(self._ctx, int(shape)))

    def tile_shape(self):
        """(shape of the last extraction, {shape: grid-pass ms of this context's last tuning})."""
        sh = ctypes.c_int32()
        ms = (ctypes.c_double * len(TILE_SHAPES))()
        self._check(lib().mdmc_get_tile_shape(self._ctx, ctypes.byref(sh), ms))
        return sh.value, {k: ms[k - 1] for k in TILE_SHAPES}

    def set_graphs(self, enable):
        """mdmc_set_graphs: replay the batched form's shape-determined front from a HIP graph."""
        self._check(lib().mdmc_set_graphs(self._ctx, 1 if enable else 0))

    def set_local_merges(self, enable):
        """mdmc_set_local_merges: plain-patch P3X3Y merges of a simplification chain update their
        twins directly (False: every merge's region is rebuilt; the mesh is the same)."""
        self._check(lib().mdmc_set_local_merges(self._ctx, 1 if enable else 0))

    def set_sequential_merges(self, enable):
        """mdmc_set_sequential_merges: the splitter's merge phase always takes the one-thread
        ascending pass it otherwise takes only on schedule conflicts (the mesh is the same)."""
        self._check(lib().mdmc_set_sequential_merges(self._ctx, 1 if enable else 0))

    def graph_replays(self):
        n = ctypes.c_int64()
        self._check(lib().mdmc_get_graph_replays(self._ctx, ctypes.byref(n)))
        return n.value

    def set_stream(self, stream_ptr):
        st = lib().mdmc_set_stream(self._ctx, ctypes.c_void_p(stream_ptr))
        self._check(st)

    def _check(self, st):
        if st != 0:
            raise MdmcError(st, lib().mdmc_last_error(self._ctx).decode())

    def extract_raw(self, data_ptr, dims, origin, delta, iso, flags, refinement=0, on_device=False):
        """Runs the pipeline on a grid given by pointer; the mesh stays on the device."""
        g = _Grid()
        g.dims[:] = [int(d) for d in dims]
        g.origin[:] = [float(np.float32(o)) for o in origin]
        g.delta[:] = [float(np.float32(d)) for d in delta]
        g.data = ctypes.c_void_p(data_ptr)
        g.data_on_device = 1 if on_device else 0
        o = _Options(int(flags), int(refinement), int(self.splitting_step), int(self.max_split_iterations))
        st = Stats()
        self.generation += 1
        self._check(lib().mdmc_extract(self._ctx, ctypes.byref(g), ctypes.c_float(iso), ctypes.byref(o), ctypes.byref(st)))
        self.last_stats = st
        self.n_problems = 1
        return st

    def extract_batch_raw(self, grids, isos_per_grid, flags, refinement=0):
        """mdmc_extract_batch.  grids: list of (data_ptr, dims (nx, ny, nz), origin, delta, on_device);
        isos_per_grid: list (one per grid) of equally long iso-value lists.  Returns total Stats."""
        G = len(grids)
        I = len(isos_per_grid[0])
        arr = (_Grid * G)()
        for i, (ptr, dims, origin, delta, on_device) in enumerate(grids):
            arr[i].dims[:] = [int(d) for d in dims]
            arr[i].origin[:] = [float(np.float32(o)) for o in origin]
            arr[i].delta[:] = [float(np.float32(d)) for d in delta]
            arr[i].data = ctypes.c_void_p(ptr)
            arr[i].data_on_device = 1 if on_device else 0
        isos = np.ascontiguousarray(np.asarray(isos_per_grid, np.float32).reshape(G * I))
        o = _Options(int(flags), int(refinement), int(self.splitting_step), int(self.max_split_iterations))
        st = Stats()
        self.generation += 1
        self._check(lib().mdmc_extract_batch(self._ctx, arr, G, isos.ctypes.data_as(ctypes.c_void_p), I, ctypes.byref(o), ctypes.byref(st)))
        self.last_stats = st
        self.n_problems = G * I
        return st

    def extract_batch(self, fields, isos_per_grid, origins, deltas, flags, refinement=0):
        """Host float32 fields [z][y][x]; returns the total Stats (meshes stay resident)."""
        keep = [np.ascontiguousarray(f, np.float32) for f in fields]
        grids = [(f.ctypes.data, (f.shape[2], f.shape[1], f.shape[0]), o, d, False) for f, o, d in zip(keep, origins, deltas)]
        return self.extract_batch_raw(grids, isos_per_grid, flags, refinement)

    def batch_offsets(self):
        """(vertex, halfedge, face) offsets, each n_problems + 1 long."""
        n = getattr(self, "n_problems", 1) + 1
        v, h, f = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int64)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self._check(lib().mdmc_batch_offsets(self._ctx, p(v), p(h), p(f)))
        return v, h, f

    def batch_stats(self):
        arr = (Stats * self.n_problems)()
        self._check(lib().mdmc_batch_stats(self._ctx, arr))
        return [arr[i] for i in range(self.n_problems)]

    def batch_meshes(self):
        """The resident batch split into per-problem HalfedgeMesh objects with their own ids."""
        m = self.mesh_numpy()
        vo, ho, fo = self.batch_offsets()
        out = []
        for b in range(len(vo) - 1):
            he = m.halfedges[ho[b]:ho[b + 1]].copy()
            if len(he):
                he[:, 0] -= vo[b]
                he[:, 1] -= fo[b]
                he[:, 2] -= ho[b]
                he[:, 3] = np.where(he[:, 3] >= 0, he[:, 3] - ho[b], he[:, 3])
            out.append(HalfedgeMesh(m.positions[vo[b]:vo[b + 1]], m.normals[vo[b]:vo[b + 1]], he, m.face_halfedge[fo[b]:fo[b + 1]] - ho[b],
                                    m.face_attr[fo[b]:fo[b + 1]], m.vertex_flags[vo[b]:vo[b + 1]], m.vertex_halfedge[vo[b]:vo[b + 1]]))
        return out

    def backward(self, dpos_ptr, dfield_ptr, accumulate=False):
        """mdmc_backward: device dL/d(vertex position) [V,3] -> device dL/d(field) of the
        last extraction, which must have used RECORD_ADJOINT."""
        self._check(lib().mdmc_backward(self._ctx, ctypes.c_void_p(dpos_ptr), ctypes.c_void_p(dfield_ptr), 1 if accumulate else 0))

    def backward_noalloc(self, dpos_ptr, dfield_ptr, accumulate=False):
        """mdmc_backward_noalloc: mdmc_backward with its scratch in the library's fixed per-device
        array (no device allocation; MdmcError with status MDMC_ERR_CAPACITY past its size)."""
        self._check(lib().mdmc_backward_noalloc(self._ctx, ctypes.c_void_p(dpos_ptr), ctypes.c_void_p(dfield_ptr), 1 if accumulate else 0))

    def copy_device(self, positions=0, normals=0, halfedges=0, face_halfedge=0, face_attr=0, vertex_flags=0, vertex_halfedge=0):
        """mdmc_mesh_copy into device buffers given by pointer (0 = skip)."""
        p = lambda a: ctypes.c_void_p(a) if a else None
        self._check(lib().mdmc_mesh_copy(self._ctx, p(positions), p(normals), p(halfedges), p(face_halfedge), p(face_attr),
                                         p(vertex_flags), p(vertex_halfedge), 1))

    def mesh_torch(self):
        """The resident mesh copied into new torch tensors on the context's device (needs torch):
        dict of positions [V,3] f32, normals [V,3] f32, halfedges [H,4] i32, face_halfedge [F] i32,
        face_attr [F] i16 (the uint16 bits), vertex_flags [V] u8, vertex_halfedge [V] i32."""
        import torch
        V, H, F = self.info()
        dev = torch.device("cuda", torch.cuda.current_device())
        t = {"positions": torch.empty((V, 3), dtype=torch.float32, device=dev),
             "normals": torch.empty((V, 3), dtype=torch.float32, device=dev),
             "halfedges": torch.empty((H, 4), dtype=torch.int32, device=dev),
             "face_halfedge": torch.empty(F, dtype=torch.int32, device=dev),
             "face_attr": torch.empty(F, dtype=torch.int16, device=dev),
             "vertex_flags": torch.empty(V, dtype=torch.uint8, device=dev),
             "vertex_halfedge": torch.empty(V, dtype=torch.int32, device=dev)}
        torch.cuda.synchronize()
        self.copy_device(**{k: (v.data_ptr() if v.numel() else 0) for k, v in t.items()})
        return t

    def generate_volume(self, dims, surface_case):
        """UniformGridHost::generateVolume(dims, surfaceCase) (UniformGrid.inl:463-486): the implicit
        surface sampled at dims = (nx, ny, nz) points of [-1, 1]^3 on the GPU.  Returns
        (values float32 [nz, ny, nx], origin, delta)."""
        nx, ny, nz = (int(d) for d in dims)
        d3 = np.array([nx, ny, nz], np.int32)
        vals = np.zeros((nz, ny, nx), np.float32)
        o = np.zeros(3, np.float32)
        dl = np.zeros(3, np.float32)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self._check(lib().mdmc_generate_surface(self._ctx, int(surface_case), p(d3), p(vals), 0, p(o), p(dl)))
        return vals, tuple(float(x) for x in o), tuple(float(x) for x in dl)

    def generate_edge_case(self, edge_case):
        """UniformGridHost::generateEdgeCase (UniformGrid.inl:489-509): (values [3, 4, 4], origin, delta)."""
        vals = np.zeros((3, 4, 4), np.float32)
        d3 = np.zeros(3, np.int32)
        o = np.zeros(3, np.float32)
        dl = np.zeros(3, np.float32)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self._check(lib().mdmc_generate_edge_case(self._ctx, int(edge_case), p(vals), 0, p(d3), p(o), p(dl)))
        assert tuple(d3) == (4, 4, 3)
        return vals, tuple(float(x) for x in o), tuple(float(x) for x in dl)

    def processed_grid(self, index=0):
        """The padded / refined grid the last extraction ran on (UniformGridHost::readFromFile,
        UniformGrid.inl:573-608), grid `index` of a batch: (values float32 [nz, ny, nx], origin, delta)."""
        d3 = np.zeros(3, np.int32)
        o = np.zeros(3, np.float32)
        dl = np.zeros(3, np.float32)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self._check(lib().mdmc_processed_grid(self._ctx, int(index), None, 0, p(d3), p(o), p(dl)))
        vals = np.zeros((int(d3[2]), int(d3[1]), int(d3[0])), np.float32)
        self._check(lib().mdmc_processed_grid(self._ctx, int(index), p(vals), 0, None, None, None))
        return vals, tuple(float(x) for x in o), tuple(float(x) for x in dl)

    def triangles(self):
        """mdmc_mesh_triangles: uint32 [F, 6], two triangles (x,y,z),(z,w,x) per face."""
        F = self.info()[2]
        idx = np.zeros((F, 6), np.uint32)
        if F:
            self._check(lib().mdmc_mesh_triangles(self._ctx, idx.ctypes.data_as(ctypes.c_void_p), 0))
        return idx

    def face_colors(self):
        """mdmc_mesh_face_colors: uint8 [F], the 5-colouring of OriginalDMC's FaceColoring."""
        F = self.info()[2]
        out = np.zeros(F, np.uint8)
        if F:
            self._check(lib().mdmc_mesh_face_colors(self._ctx, out.ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def triangulate_maxmin(self):
        """mdmc_mesh_triangulate_maxmin: (uint32 [F, 6] triangles, float32 [F, 2] mean ratios)."""
        F = self.info()[2]
        tri = np.zeros((F, 6), np.uint32)
        q = np.zeros((F, 2), np.float32)
        if F:
            self._check(lib().mdmc_mesh_triangulate_maxmin(self._ctx, tri.ctypes.data_as(ctypes.c_void_p), q.ctypes.data_as(ctypes.c_void_p), 0))
        return tri, q

    def render_vertices(self):
        """mdmc_mesh_render_vertices: float32 [V, 10] = position, normal, rgba by vertex flags."""
        V = self.info()[0]
        out = np.zeros((V, 10), np.float32)
        if V:
            self._check(lib().mdmc_mesh_render_vertices(self._ctx, out.ctypes.data_as(ctypes.c_void_p), 0))
        return out

    def info(self):
        i = _Info()
        self._check(lib().mdmc_mesh_get_info(self._ctx, ctypes.byref(i)))
        return i.n_vertices, i.n_halfedges, i.n_faces

    def mesh_numpy(self):
        V, H, F = self.info()
        pos = np.zeros((V, 3), np.float32)
        nrm = np.zeros((V, 3), np.float32)
        he = np.zeros((H, 4), np.int32)
        fh = np.zeros(F, np.int32)
        fa = np.zeros(F, np.uint16)
        vf = np.zeros(V, np.uint8)
        vh = np.zeros(V, np.int32)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        self._check(lib().mdmc_mesh_copy(self._ctx, p(pos), p(nrm), p(he), p(fh), p(fa), p(vf), p(vh), 0))
        return HalfedgeMesh(pos, nrm, he, fh, fa, vf, vh)

    def run(self, iso, values, origin=(0.0, 0.0, 0.0), delta=(1.0, 1.0, 1.0), refinement=0, add_zero_boundary=True):
        """DualMarchingCubes::run(iso, mesh, grid): values is a float32 array [z][y][x] on the host."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        nz, ny, nx = v.shape
        self.extract_raw(v.ctypes.data, (nx, ny, nz), origin, delta, iso, self.flags(add_zero_boundary), refinement, False)
        return self.mesh_numpy()

    # statistics getters of DualMarchingCubes (DualMarchingCubes.h:98-118)
    def getNumVertices(self):
        return self.last_stats.n_dmc_vertices

    def getNumQuads(self):
        return self.last_stats.n_dmc_quads

    def getNumNonManifoldEdges4(self):
        return self.last_stats.nm_edges4

    def getNumNonManifoldEdges3(self):
        return self.last_stats.nm_edges3

    def getNumEdges(self):
        return self.last_stats.n_edges

    def getNumNonManifoldVertices(self):
        return self.last_stats.nm_vertices

    def getNumQuadsWithNonManifoldEdges(self, k):
        return self.last_stats.quads_with_nm[k]
