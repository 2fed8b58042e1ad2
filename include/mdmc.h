/*
 * mdmc.h -- C ABI of the MI355X (gfx950) Manifold Dual Marching Cubes library.
 *
 * Drop-in for the DMC plugin path of PhiliGuertler/ManifoldDMC:
 *   DMC::MarchingCubes::run(float isoValue, Mesh&, const UniformGridHost<float>&)
 *     (MarchingCubes/VisualDMC/src/DMC/DualMarchingCubes.h:58), implemented by
 *   DMC::DualMarchingCubes::run (DualMarchingCubes.cu:220-289) with
 *   runPostProcesses (DualMarchingCubes.cu:335-380).
 * The grid flags of UniformGridHost::readFromFile(file, refinement, addZeroBoundary)
 * (UniformGrid.h:388, UniformGrid.inl:573-608) and the boolean members of
 * DualMarchingCubes (DualMarchingCubes.cu:110-133) become mdmc_options.
 *
 * Plain C types only.  Ids are int32 like the reference's IndexType.  All
 * device pointers are HIP device pointers on the context's device.
 */
#ifndef MDMC_H
#define MDMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference has none: it logs, throws or truncates) ---- */
typedef enum {
    MDMC_OK = 0,
    MDMC_ERR_INVALID_ARG = 1,
    MDMC_ERR_OOM = 2,
    MDMC_ERR_CAPACITY = 3,   /* a count exceeds the int32 id range */
    MDMC_ERR_HASH_FULL = 4,  /* reserved: the library sizes every table from exact counts */
    MDMC_ERR_HIP = 5,
    MDMC_ERR_NO_RESULT = 6,
    MDMC_ERR_INTERNAL = 7    /* an id produced by a stage is out of range; the run was stopped */
} mdmc_status;

/* ---- option bits (DualMarchingCubes members, DualMarchingCubes.cu:110-133;
 *      addZeroBoundary of UniformGridHost::readFromFile, UniformGrid.inl:575) ---- */
#define MDMC_PAD1            0x01 /* pad by one clamped (duplicated) layer, UniformGrid.inl:290-295 */
#define MDMC_SURF_ITER       0x02 /* m_enableSurfaceToVertexIteration (default on) */
#define MDMC_P3X3Y_COLOR     0x04 /* m_simplifyP3x3yColor, HalfedgeMesh::simplify3X3Y */
#define MDMC_P3X3Y_OLD       0x08 /* m_simplifyP3x3yOld, HalfedgeMesh::simplify3X3YOld */
#define MDMC_P3333           0x10 /* m_simplifyP3333, HalfedgeMesh::simplify3333 */
#define MDMC_SPLIT_NM        0x20 /* m_splitNonManifoldHalfedges, ManifoldSplitter::run */
#define MDMC_DETECT_NM_VERTS 0x40 /* m_detectNonManifoldVertices */
#define MDMC_RECORD_ADJOINT  0x80 /* no reference counterpart: keep the DMC topology for mdmc_backward
                                     (needs MDMC_SURF_ITER, P3X3Y*, P3333 and SPLIT_NM off) */
#define MDMC_DEFAULT_FLAGS   (MDMC_PAD1 | MDMC_SURF_ITER | MDMC_SPLIT_NM | MDMC_DETECT_NM_VERTS)
#define MDMC_ALL_POST        (MDMC_P3X3Y_COLOR | MDMC_P3X3Y_OLD | MDMC_P3333 | MDMC_SPLIT_NM | MDMC_DETECT_NM_VERTS)

/* SplittingStep, ManifoldSplitter.h:155-166 */
#define MDMC_SPLIT_STEP_NONE 0
#define MDMC_SPLIT_STEP_SUBDIVIDE 1
#define MDMC_SPLIT_STEP_COLLAPSE 2
#define MDMC_SPLIT_STEP_MERGE 3
#define MDMC_SPLIT_STEP_TWINS 4
#define MDMC_SPLIT_STEP_MOVE 5
#define MDMC_SPLIT_STEP_ITERATIVE 6
#define MDMC_SPLIT_STEP_ALL 100

typedef struct {
    int32_t flags;           /* MDMC_* bits */
    int32_t refinement;      /* refinement cuts per cell (UniformGrid.inl:587), >= 0 */
    int32_t split_step;      /* MDMC_SPLIT_STEP_* (reference default: ALL) */
    int32_t max_split_iters; /* guard for the reference's unbounded split loop; <=0 means 64 */
} mdmc_options;

/* A uniform scalar grid, x fastest then y then z (UniformGrid.h:50-52).
 * origin/delta as in UniformGridStats (UniformGrid.h:121-129). */
typedef struct {
    int32_t dims[3];
    float origin[3];
    float delta[3];
    const float* data;
    int32_t data_on_device; /* 1: data is a device pointer, 0: host pointer */
} mdmc_grid;

/* Counters of the last extraction (LayerCUDA.cpp:246-255 statistics; HalfedgeMesh.cu:1164-1239). */
typedef struct {
    int64_t n_vertices, n_halfedges, n_faces;
    int64_t n_dmc_vertices, n_dmc_quads;
    int64_t nm_edges4, nm_edges3, n_edges, nm_vertices; /* nm_vertices = -1 when detection is off */
    int64_t quads_with_nm[5];                           /* faces with k edges of >2 faces, k = 0..4 */
    int64_t split_iterations;
    int64_t anomalies; /* situations the reference leaves undefined (out-of-range ids it would
                          read, fan walks past 4096 steps, buckets past 16 halfedges); 0 normally */
    int64_t cells;     /* cells of the processed grid */
    /* stage times (ms, HIP events on the context's stream): whole call, pad/refine, DMC grid
     * pass, halfedge build, simplifications, twin repair, split, detection */
    double ms_total, ms_grid, ms_dmc, ms_halfedge, ms_simplify, ms_fix_twins, ms_split, ms_detect;
    /* the grid-pass kernel alone (classification, dual vertices, quads) */
    double ms_grid_pass_kernel;
    /* split iterations whose merge phase took the one-thread ascending pass (the parallel merge
     * found a schedule conflict, or mdmc_set_sequential_merges forced it); -1 for batches */
    int64_t split_sequential_merges;
    /* tiles of the single-pass prefix sums and compactions whose wait on the tiles before them
     * passed its spin limit and that summed their prefix directly instead (exact either way;
     * 0 normally); -1 in a batch's per-problem statistics, whose call statistics hold the count */
    int64_t lookback_fallbacks;
} mdmc_stats;

/* Sizes of the resident result. */
typedef struct {
    int64_t n_vertices, n_halfedges, n_faces;
} mdmc_mesh_info;

typedef struct mdmc_ctx mdmc_ctx;

/* One context per host thread; it owns a HIP stream and reuses its buffers across
 * calls (the reference keeps 20M-capacity buffers in the DualMarchingCubes object,
 * DualMarchingCubes.cu:291-313). device < 0 means the current device. */
mdmc_status mdmc_create(mdmc_ctx** out, int32_t device);
void mdmc_destroy(mdmc_ctx* ctx);
const char* mdmc_last_error(const mdmc_ctx* ctx);

/* Run all following work on this stream (hipStream_t passed as void*); NULL = own stream. */
mdmc_status mdmc_set_stream(mdmc_ctx* ctx, void* hip_stream);

/* Shape of the grid pass (no reference counterpart; the reference launches one thread
 * per grid point, DualMarchingCubes.cu:38-67).  MDMC_TILE_AUTO (the default): on the first
 * extraction of a workload class (cell count and problem count to the next power of two,
 * move-to-surface on/off) the context times every shape on that input and keeps the
 * fastest for the class.  Any other value forces that shape.  Every shape gives the same
 * mesh.  There is no hash table to tune: edges and halfedges are matched through exact-size
 * lists, not open addressing. */
#define MDMC_TILE_AUTO 0
#define MDMC_TILE_1024 1          /* 1024 cells per tile, tiles dealt to XCDs in contiguous runs */
#define MDMC_TILE_2048 2
#define MDMC_TILE_4096 3
#define MDMC_TILE_1024_LINEAR 4   /* the same tiles in launch order */
#define MDMC_TILE_2048_LINEAR 5
#define MDMC_TILE_4096_LINEAR 6
#define MDMC_TILE_SLAB4 7         /* row slabs: a block owns up to 254 cells of 4 rows and marches through
                                     a chunk of slices; counts per row segment, one scan, then the emit */
#define MDMC_TILE_SLAB2 8         /* the same with 2 rows */
#define MDMC_TILE_2048_QUEUE 9    /* 2048-cell tiles taken from a work queue by persistent blocks (as many
                                     as the GPU holds at once), the tables loaded once per block */
#define MDMC_TILE_SHAPES 9        /* number of shapes */
mdmc_status mdmc_set_tile_shape(mdmc_ctx* ctx, int32_t shape);
/* The shape the last extraction used; tune_ms (may be NULL, double[MDMC_TILE_SHAPES]) receives
 * the grid-pass times of shapes 1..MDMC_TILE_SHAPES from this context's last tuning (0 where none
 * ran). */
mdmc_status mdmc_get_tile_shape(const mdmc_ctx* ctx, int32_t* last_shape, double* tune_ms);

/* Replaces DualMarchingCubes::run(iso, mesh, grid) + runPostProcesses
 * (DualMarchingCubes.cu:220-289, :335-380).  The mesh stays on the device in the
 * context until the next call.  stats may be NULL.  Synchronises the stream. */
mdmc_status mdmc_extract(mdmc_ctx* ctx, const mdmc_grid* grid, float iso, const mdmc_options* opt, mdmc_stats* stats);

mdmc_status mdmc_mesh_get_info(const mdmc_ctx* ctx, mdmc_mesh_info* info);

/* Copies the resident mesh into caller buffers (any may be NULL):
 *   positions/normals float[3*nV], halfedges int32[4*nH] = {origin, face, next, twin}
 *   (Halfedge.h:120-210), face_halfedge int32[nF], face_attr uint16[nF]
 *   (QuadrilateralAttribute bits, DeviceVectors.h:334-450), vertex_flags uint8[nV]
 *   (Flags, Halfedge.h:14-111), vertex_halfedge int32[nV].
 * dst_on_device: 1 = device buffers, 0 = host buffers. */
mdmc_status mdmc_mesh_copy(const mdmc_ctx* ctx, float* positions, float* normals, int32_t* halfedges,
                           int32_t* face_halfedge, uint16_t* face_attr, uint8_t* vertex_flags,
                           int32_t* vertex_halfedge, int32_t dst_on_device);

/* Batched form (no reference counterpart: the reference extracts one grid per call).
 * n_grids grids, each at isos_per_grid iso-values (isos[g * isos_per_grid + k]), form
 * n_grids * isos_per_grid problems, problem b = g * isos_per_grid + k, all with the same
 * options.  Every problem gets exactly the mesh mdmc_extract gives it; the meshes are
 * resident CSR-packed: problem b owns vertices [vertex_offsets[b], vertex_offsets[b+1]),
 * halfedges and faces likewise, and the ids in halfedges / face_halfedge are indices into
 * the concatenation (subtract the problem's offsets for its own ids).  vertex_halfedge
 * holds each problem's own values.  All problems run through one grid-pass launch and
 * one pipeline.  stats (may be NULL) holds totals.  Synchronises the stream. */
mdmc_status mdmc_extract_batch(mdmc_ctx* ctx, const mdmc_grid* grids, int32_t n_grids, const float* isos, int32_t isos_per_grid,
                               const mdmc_options* opt, mdmc_stats* stats);
/* Graph replay of the batched form (no reference counterpart).  For a batch whose grids are
 * all on the device, once its tile shape is chosen, the part of the pipeline that the grid
 * shapes alone determine -- the descriptor uploads, padding / refinement, classification and
 * the tile-count scans -- is captured into a HIP graph on the second call with the same
 * shapes, options and buffers, and replayed (with the new calls' grid pointers and iso-values)
 * on every later one; the stages after it size their buffers from the counts and run eagerly.
 * Results are the same with or without.  enable = 0 turns it off and releases the graph
 * (default on).  mdmc_get_graph_replays: how many extractions of this context ran the graph. */
mdmc_status mdmc_set_graphs(mdmc_ctx* ctx, int32_t enable);
mdmc_status mdmc_get_graph_replays(const mdmc_ctx* ctx, int64_t* replays);
/* Local twin update in a chain of simplifications (no reference counterpart).  While another
 * simplification follows, a P3X3Y merge whose neighbourhood is a plain patch rewrites its two
 * new twin pairs directly instead of having its region's buckets rebuilt; the mesh is the same
 * either way.  enable = 0 sends every merge through the region rebuild (default on). */
mdmc_status mdmc_set_local_merges(mdmc_ctx* ctx, int32_t enable);
/* Verification switch for the splitter's merge phase (no reference counterpart).  The merges
 * of ManifoldSplitter's mergeHalfedges (ManifoldSplitter.cu:1021-1067) run in parallel and fall
 * back to a one-thread pass in ascending order only when two of them conflict; enable = 1
 * makes every iteration take that pass, so tests can compare it with the parallel one and
 * with the reference's results (default 0).  The mesh is the same either way. */
mdmc_status mdmc_set_sequential_merges(mdmc_ctx* ctx, int32_t enable);

/* Per-problem offsets of the resident result (host arrays of n_problems + 1; any may be NULL).
 * After mdmc_extract (one problem) they are {0, n}. */
mdmc_status mdmc_batch_offsets(const mdmc_ctx* ctx, int64_t* vertex_offsets, int64_t* halfedge_offsets, int64_t* face_offsets);
/* Per-problem counts of the last mdmc_extract_batch (host array of n_problems; times are 0,
 * anomalies and cells -1). */
mdmc_status mdmc_batch_stats(const mdmc_ctx* ctx, mdmc_stats* per_problem);

/* Adjoint of the last mdmc_extract (no reference counterpart; the reference has no backward
 * pass).  That extract must have had MDMC_RECORD_ADJOINT.  With the topology frozen, maps
 * dL_dpos (device float[3 * n_vertices], d loss / d vertex position, the mesh's vertex order)
 * to dL_dfield (device float[dims[0]*dims[1]*dims[2]] of the grid passed to mdmc_extract),
 * through the dual-vertex formula of CellIntersection::sliceP (CellIntersection.inl:368-439)
 * and the padding / refinement of refineGrid (UniformGrid.inl:279-319).  accumulate = 1 adds
 * to dL_dfield instead of overwriting it.  Asynchronous on the context's stream; deterministic. */
mdmc_status mdmc_backward(mdmc_ctx* ctx, const float* dL_dpos, float* dL_dfield, int32_t accumulate);

/* No reference counterpart: mdmc_backward for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and the same result bit for bit; it calls no hipMalloc / hipFree and no allocating
 * routine.  Its scratch (8 floats per dual vertex, plus one per point of the padded / refined
 * grid when padding or refinement was used) is a fixed 32 Mi-float (128 MiB) array per device,
 * held until the process ends; larger problems get MDMC_ERR_CAPACITY.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (several streams or contexts at once)
 * race, so serialise them.  Asynchronous on the context's stream. */
mdmc_status mdmc_backward_noalloc(mdmc_ctx* ctx, const float* dL_dpos, float* dL_dfield, int32_t accumulate);

/* Triangle index buffer, two triangles per face (x,y,z),(z,w,x): HalfedgeMesh.cu:690-714.
 * indices uint32[6*nF], device or host. */
mdmc_status mdmc_mesh_triangles(const mdmc_ctx* ctx, uint32_t* indices, int32_t dst_on_device);

/* Render vertex buffer: HalfedgeMesh insertVertices (HalfedgeMesh.cu:716-760) into the GL
 * vertex layout of DMC::Vertex (Mesh.h:15-19): per vertex 10 floats = position xyz, normal
 * xyz, colour rgba.  The colour is HSV(hue, 0.8, 0.8) -> RGB (LazyEngine Color::HSVtoRGB)
 * with hue 180 (cyan) by default, 0 on a non-manifold edge, 60 on an edge of 3 faces, 270 for
 * a vertex with several fans, 300 for non-manifold + 3-face, 240 for non-manifold + several
 * fans.  A vertex without an outgoing halfedge (left unreferenced by a merge) stays all
 * zeros, as the reference's buffer and the .obj files written from it (Mesh.cu:446-506) show.
 * vertices float[10*nV], device or host. */
mdmc_status mdmc_mesh_render_vertices(const mdmc_ctx* ctx, float* vertices, int32_t dst_on_device);

/* 5-colouring of the faces (OriginalDMC FaceColoring::colorFaces, FaceColoring.cu:5-141): the
 * faces' grid-edge colours (0..23, face_attr bits 0-4) with 5..23 replaced class by class by the
 * first of 0..4 no neighbour has, then colour-4 faces moved to a free colour of 0..3.  Faces
 * without a grid colour (31, e.g. after a simplification) keep 31.  The resident mesh is not
 * changed.  colors uint8[nF], device or host. */
mdmc_status mdmc_mesh_face_colors(const mdmc_ctx* ctx, uint8_t* colors, int32_t dst_on_device);
/* Quad -> triangle split by the MaxMin angle criterion with the mean-ratio quality of every
 * triangle (OriginalDMC DualMarchingCubes.cu:51-107, Triangles.h:80-90, QualityMeasure.h:37-70):
 * per face (v0..v3) either (v1,v3,v0),(v1,v2,v3) or (v0,v1,v2),(v0,v2,v3), whichever cut has
 * the larger smallest angle; quality[2f + k] = 2 sqrt(3) |cross| / (sum of squared edge lengths)
 * of triangle k (1 = equilateral).  indices uint32[6*nF], quality float[2*nF] (may be NULL). */
mdmc_status mdmc_mesh_triangulate_maxmin(const mdmc_ctx* ctx, uint32_t* indices, float* quality, int32_t dst_on_device);

/* ---- synthetic grids (UniformGridHost::generateVolume / generateEdgeCase) ---- */
/* SurfaceCase, UniformGrid.h:26-37; the surfaces of UniformGrid.inl:25-127 */
#define MDMC_SURFACE_SPHERE 0
#define MDMC_SURFACE_TORUS 1
#define MDMC_SURFACE_TWO_HOLED_TORUS 2
#define MDMC_SURFACE_FOUR_HOLED_TORUS 3
#define MDMC_SURFACE_GENUS_TWO 4
#define MDMC_SURFACE_IWP 5
#define MDMC_SURFACE_PW_HYBRID 6
#define MDMC_SURFACE_NEOVIUS 7
#define MDMC_SURFACE_GOURSAT 8
#define MDMC_SURFACE_STEINER_ROMAN 9
/* EdgeCase, UniformGrid.h:39-42 */
#define MDMC_EDGE_NM_VERTEX0 0
#define MDMC_EDGE_NM_VERTEX1 1

/* Replaces UniformGridHost::generateVolume(dims, surfaceCase) (UniformGrid.inl:463-486,
 * fillVolume :137-189): dims[0]*dims[1]*dims[2] values (x fastest) of the implicit surface at
 * the points origin + index * delta of [-1, 1]^3, delta = 2 / (dims - 1).  Iso-value 0 gives
 * the surface.  values: device or host float array; origin / delta (may be NULL) receive the
 * grid geometry.  Synchronises the context's stream. */
mdmc_status mdmc_generate_surface(mdmc_ctx* ctx, int32_t surface_case, const int32_t dims[3], float* values, int32_t dst_on_device,
                                  float origin[3], float delta[3]);
/* Replaces UniformGridHost::generateEdgeCase (UniformGrid.inl:489-509, fillEdgeCase :191-269):
 * the 4x4x3 grid (48 values, 850 / 950) with a non-manifold vertex at iso 900.  dims / origin /
 * delta (may be NULL) receive {4,4,3}, -1, 2 / (dims - 1). */
mdmc_status mdmc_generate_edge_case(mdmc_ctx* ctx, int32_t edge_case, float* values, int32_t dst_on_device, int32_t dims[3], float origin[3],
                                    float delta[3]);

/* Replaces reading back the grid that UniformGridHost::readFromFile(file, refinement,
 * addZeroBoundary) builds (UniformGrid.inl:573-608; getDimensions / getDeltas / getOrigin,
 * UniformGrid.h:316-345): the padded / refined grid the last extraction ran on, grid
 * grid_index of a batch.  values (may be NULL to query the geometry only) receives
 * dims[0]*dims[1]*dims[2] floats, x fastest, on the device or the host; dims / origin / delta
 * (may be NULL) receive the geometry.  Synchronises the context's stream. */
mdmc_status mdmc_processed_grid(mdmc_ctx* ctx, int32_t grid_index, float* values, int32_t dst_on_device, int32_t dims[3], float origin[3],
                                float delta[3]);

#ifdef __cplusplus
}
#endif

#endif /* MDMC_H */
