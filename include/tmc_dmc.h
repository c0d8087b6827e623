/*
 * tmc_dmc.h -- C ABI of the MI355X Dual Marching Cubes engine (libtmc_dmc.so).
 *
 * This is the drop-in boundary for the reference's extraction path.  The
 * reference has no library boundary of its own; its two call surfaces are
 *   cpp_dmc/DualMarchingCubes.h:163-165
 *     void dmc::DualMarchingCubes::dualMC(const double i0, UGrid& ugrid,
 *          std::vector<Vertex>& v, std::vector<Normal>& n,
 *          std::vector<int>& tris, std::vector<int>& quads, const bool sFlag);
 *   dmc/DualMarchingCubes.h:74-77
 *     void p_mc::DualMarchingCubes::dualMC(const float i0, v, n, t, q,
 *          he, hef, hev, std::map<std::string,int>& config);
 * plus the grid set-up of cpp_dmc/UniformGrid.h (init/scalar/gradient) and
 * getBndVertices (cpp_dmc/DualMarchingCubes.h:171-182).  Each entry point
 * below names the reference interface it replaces.  include/tmc_dmc.hpp
 * rebuilds the reference's C++ classes on top of this ABI.
 *
 * All entry points return a dmc_status; none of them exits the process
 * (the reference exits on CUDA and file errors, dmc/helper_cuda.h:88-98,
 * cpp_dmc/UniformGrid.cpp:7-9).  A context owns one HIP stream and its
 * device buffers; calls on one context must not overlap.
 */
#ifndef TMC_DMC_H
#define TMC_DMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DMC_OK = 0,
    DMC_ERR_INVALID_ARG = 1,   /* bad pointer, size or flag */
    DMC_ERR_OUT_OF_MEMORY = 2, /* device allocation failed */
    DMC_ERR_CAPACITY = 3,      /* a count exceeds the 32-bit index range */
    DMC_ERR_HIP = 4,           /* a HIP runtime call failed */
    DMC_ERR_NO_DEVICE = 5,     /* no usable GPU */
    DMC_ERR_BUDGET = 6,        /* a request does not fit the memory budget */
    DMC_ERR_STATE = 7          /* call out of order (e.g. download before extract) */
} dmc_status;

typedef struct dmc_context dmc_context;

/* Uniform grid description: nx*ny*nz points, x fastest, point (i,j,k) at
 * (x0 + i*dx, y0 + j*dy, z0 + k*dz).  Replaces the geometry part of
 * dmc::UniformGrid (cpp_dmc/UniformGrid.h:18-354). */
typedef struct {
    int32_t nx, ny, nz;
    double x0, y0, z0;
    double dx, dy, dz;
} dmc_grid;

/* field element types */
#define DMC_FIELD_F32 0
#define DMC_FIELD_F64 1

/* extraction flags (the CUDA reference's config keys, dmc/kernel.cu:25-32) */
#define DMC_SIMPLIFY_3X3Y_COLOR    (1u << 0)  /* "p3X3YColor" */
#define DMC_SIMPLIFY_3X3Y_ISOLATED (1u << 1)  /* "p3X3YOld"   */
#define DMC_SIMPLIFY_3333          (1u << 2)  /* "p3333"      */
#define DMC_TRIANGLES              (1u << 3)  /* collectTriangles */
#define DMC_GRADIENTS_SUPPLIED     (1u << 4)  /* normals from caller gradients */
#define DMC_FIELD_ON_DEVICE        (1u << 5)  /* field/gradients are device pointers */
#define DMC_STAGE_TIMING           (1u << 6)  /* record per-stage HIP events */
/* cpp sFlag == true (cpp_dmc/DualMarchingCubes.cpp:62-76) */
#define DMC_SIMPLIFY_ALL (DMC_SIMPLIFY_3X3Y_COLOR | DMC_SIMPLIFY_3X3Y_ISOLATED | DMC_SIMPLIFY_3333)

/* Result of one extraction.  Device pointers are owned by the context and
 * stay valid until the next dmc_extract or dmc_destroy on it. */
typedef struct {
    int64_t nv;                 /* vertices */
    int64_t nq;                 /* quadrilaterals */
    int64_t nt;                 /* triangles (DMC_TRIANGLES) */
    const double* d_positions;  /* 3*nv, FP64 */
    const double* d_normals;    /* 3*nv, FP64, unit length */
    const int32_t* d_quads;     /* 4*nq vertex ids */
    const int32_t* d_tris;      /* 3*nt vertex ids */
    const int32_t* d_twins;     /* 4*nq: twin of halfedge 4q+k, -1 on the boundary */
    const uint8_t* d_he_flags;  /* 4*nq: 0 manifold, 1 non-manifold, 2 multiple boundary edge */
    const int32_t* d_he_v;      /* nv: a halfedge leaving each vertex */
    int64_t nonmanifold_edges;  /* edges with 3 or 4 faces in the final mesh */
    int64_t nbnd;               /* entries of getBndVertices() */
    int64_t edge_errors;        /* edges with more than 4 faces (left unpaired) */
    int64_t invalid_refs;       /* quad vertex ids out of range after simplification */
    int64_t valence_zero;       /* vertices without faces at the last valence pass */
    uint32_t failed_projections_l1;  /* nrProj1 (cpp_dmc/DualMarchingCubes.cpp:51) */
    uint32_t failed_projections_l2;  /* nrProj2 (:52) */
    int64_t active_cells;       /* cells that are neither all-in nor all-out */
    int64_t nv_emitted;         /* vertices produced by the cell sweep */
    int64_t nv_cells;           /* vertices after removeUnusedVertices */
    int64_t nq_cells;           /* quadrilaterals before simplification */
    int64_t host_syncs;         /* times the extraction waited for the device (batch-wide) */
    int64_t pass_mode;          /* how its pass ran: 0 launch by launch, 1 captured into a HIP graph and
                                   launched, 2 a replay of the graph of an earlier pass of the same shape */
} dmc_mesh;

/* Create a context on HIP device `device`. */
dmc_status dmc_create(dmc_context** ctx, int device);
/* Release a context and all of its device memory. */
dmc_status dmc_destroy(dmc_context* ctx);

/* Extract the iso-surface of `field` at `iso`.  Replaces
 * cpp_dmc/DualMarchingCubes.h:163-165 (dualMC, sFlag -> DMC_SIMPLIFY_ALL)
 * and dmc/DualMarchingCubes.h:74-77 (dualMC with its config map).
 * field_type is DMC_FIELD_F32 or DMC_FIELD_F64.  Without
 * DMC_GRADIENTS_SUPPLIED the normals use the gradient of
 * UniformGrid::estimateGradient + flip_gradient (cpp_dmc/UniformGrid.cpp:
 * 156-197), computed on the fly; with it, `gradients` holds 3 doubles per
 * grid point as cpp_dmc's UniformGrid stores them.  The mesh stays on the
 * device; `out` receives counts and device pointers. */
dmc_status dmc_extract(dmc_context* ctx, const dmc_grid* grid, const void* field, int field_type,
                       double iso, uint32_t flags, const double* gradients, dmc_mesh* out);

/* One volume of a batch: grid, field (element type shared by the batch),
 * optional gradients (DMC_GRADIENTS_SUPPLIED) and iso-value.  With
 * DMC_FIELD_ON_DEVICE the pointers are device pointers. */
typedef struct {
    dmc_grid grid;
    const void* field;
    const double* gradients;
    double iso;
} dmc_volume;

/* Extract many independent volumes in one pass (the batched mode; no
 * counterpart in the reference, which extracts one grid per dualMC call,
 * cpp_dmc/DualMarchingCubes.h:163-165).  All volumes share field_type and
 * flags.  out[v] receives volume v's mesh exactly as dmc_extract would
 * return it for that volume alone: same counts, same arrays, ids starting
 * at 0; device pointers stay valid until the next extraction on ctx.
 * Diagnostics invalid_refs and valence_zero are batch-wide; all other
 * fields are per volume. */
dmc_status dmc_extract_batch(dmc_context* ctx, const dmc_volume* vols, int32_t nvol, int field_type, uint32_t flags,
                             dmc_mesh* out);

/* ------------------------------------------------------------------------
 * Extraction that allocates nothing.  No counterpart in the reference
 * (whose dualMC grows std::vectors and, in dmc/, cudaMallocs its tables on
 * every call); it is for callers that cannot allocate device memory:
 * tight serving loops and harnesses that forbid hipMalloc.  It runs the same kernels as dmc_extract on one volume, with
 * every working buffer carved from a fixed scratch array that the library
 * holds in device memory on each GPU it runs on (64 MiB, until the process
 * ends), and writes the mesh into the caller's device buffers.
 *
 * Limits: at most 2,097,152 grid points (128^3); the working set of the
 * field's surface must fit the 64 MiB scratch (a smooth surface on 128^3
 * does, about 40 MiB; dense noise fits up to about 64^3).  Beyond either,
 * DMC_ERR_CAPACITY with the reason in dmc_last_error.  The field (and
 * gradients, DMC_GRADIENTS_SUPPLIED) must be device pointers; the flag
 * DMC_FIELD_ON_DEVICE is implied and DMC_STAGE_TIMING is ignored.
 * Calls on one GPU share its scratch array and are serialised inside the
 * library (each returns with its work finished); the array keeps the last
 * call's contents, and nothing in it is assumed zero.  The context's own
 * buffer pools are neither used nor grown; after the call, downloads and
 * diagnostics of the context return DMC_ERR_STATE until the next
 * dmc_extract.
 * ---------------------------------------------------------------------- */
typedef struct {
    double* positions;   /* device, 3 * cap_v doubles */
    double* normals;     /* device, 3 * cap_v doubles */
    int32_t* quads;      /* device, 4 * cap_q */
    int32_t* tris;       /* device, 6 * cap_q with DMC_TRIANGLES, else may be NULL */
    int32_t* twins;      /* device, 4 * cap_q: twin of halfedge 4q+k, -1 on the boundary */
    uint8_t* he_flags;   /* device, 4 * cap_q */
    int32_t* he_v;       /* device, cap_v */
    int32_t* bnd;        /* host, cap_bnd entries of getBndVertices(); may be NULL */
    int64_t cap_v, cap_q, cap_bnd;
} dmc_out_buffers;

/* As dmc_extract for one device-resident volume, allocating nothing.  `out`
 * receives the counts and the caller's buffer pointers; when the mesh does
 * not fit cap_v / cap_q (or the boundary list cap_bnd), DMC_ERR_CAPACITY
 * with out->nv / out->nq (out->nbnd) set to the sizes required. */
dmc_status dmc_extract_noalloc(dmc_context* ctx, const dmc_grid* grid, const void* field, int field_type,
                               double iso, uint32_t flags, const double* gradients,
                               const dmc_out_buffers* out_buffers, dmc_mesh* out);

/* Device bytes a context holds in its buffer pools (grown to its largest
 * extraction and kept for reuse until dmc_destroy).  No counterpart in the
 * reference; the serving loop counts it in its memory budget. */
dmc_status dmc_context_bytes(dmc_context* ctx, size_t* bytes);

/* Free a context's buffer pools and captured passes (its last mesh goes with
 * them); the next extraction allocates what it needs.  No counterpart in the
 * reference; the serving loop calls it when the worker contexts' pools would
 * keep it over its memory budget. */
dmc_status dmc_context_trim(dmc_context* ctx);

/* Copy the last mesh to host memory; any pointer may be NULL.
 * positions/normals: 3*nv doubles; quads 4*nq; tris 3*nt;
 * halfedges: 5 ints per halfedge {origin, face, next, twin, flag} as in
 * cpp_dmc/DualMarchingCubes.h:115-120 and dmc/DualMarchingCubes.cu:1098-1119;
 * he_v nv; he_f nq; bnd nbnd (cpp_dmc/DualMarchingCubes.h:171-182). */
dmc_status dmc_mesh_download(dmc_context* ctx, double* positions, double* normals, int32_t* quads,
                             int32_t* tris, int32_t* halfedges, int32_t* he_v, int32_t* he_f,
                             int32_t* bnd);
/* The same for volume `index` of the last dmc_extract_batch. */
dmc_status dmc_batch_download(dmc_context* ctx, int32_t index, double* positions, double* normals, int32_t* quads,
                              int32_t* tris, int32_t* halfedges, int32_t* he_v, int32_t* he_f, int32_t* bnd);

/* ------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, which extracts one grid per
 * synchronous dualMC call).  A server owns `workers` worker threads, each
 * with its own context (and so its own stream): a free worker admits the
 * waiting requests oldest first while their estimated working set fits
 * budget_bytes beside what the server holds -- result blocks (in use or
 * idle; idle ones are freed when that makes a request fit), the buffer
 * pools the worker contexts keep (dmc_context_bytes; a worker frees its
 * pools with dmc_context_trim before a pass while they keep the server over
 * the budget) and the estimates of the passes running -- and at most
 * max_batch volumes and max_batch_points
 * grid points per pass (0: 2^28, which keeps any pass inside the engine's
 * index limits); the first request of a pass is always admitted.  It runs
 * them as one dmc_extract_batch and completes them together; a pass that
 * fails on a size limit of the union or on one bad member is run again one
 * request at a time, so only a request that fails alone fails.  The working
 * set of a request is an estimate (lists, counts and 40 bytes per grid
 * point), so the budget bounds device memory up to that estimate.  Passes
 * of different workers overlap on the GPU, so one fills the other's host
 * round trips.  Results stay in device memory (pooled blocks) until
 * dmc_server_release.  With max_batch = 1 and one worker the same loop
 * serves one request at a time.
 * ---------------------------------------------------------------------- */
typedef struct dmc_server dmc_server;

typedef struct {
    int device;
    int field_type;          /* DMC_FIELD_F32 / DMC_FIELD_F64, for every request */
    uint32_t flags;          /* extraction flags for every request (DMC_FIELD_ON_DEVICE: device fields) */
    int32_t max_batch;       /* volumes per pass, >= 1 */
    size_t budget_bytes;     /* admission budget; 0 = 96 GiB */
    int64_t max_batch_points;  /* grid points per pass (bounds a pass's duration); 0 = 2^28 */
    int32_t workers;         /* passes run concurrently, each with its own context and stream; 0 = 1 */
} dmc_server_opts;

typedef struct {
    dmc_mesh mesh;           /* counts and device pointers (valid until release) */
    double t_submit;         /* seconds since dmc_server_create */
    double t_start;          /* its batch started */
    double t_done;           /* its batch completed */
    int32_t batch_size;      /* volumes in that batch */
    dmc_status status;
} dmc_result;

dmc_status dmc_server_create(dmc_server** server, const dmc_server_opts* opts);
/* Finishes the queued requests, then frees everything. */
dmc_status dmc_server_destroy(dmc_server* server);
/* Queue one volume; its field must stay valid until the request is done.
 * A grid the engine cannot take (an axis below 2 or above 2^24 points,
 * more than 2^32 cells) is refused here with DMC_ERR_INVALID_ARG or
 * DMC_ERR_CAPACITY. */
dmc_status dmc_server_submit(dmc_server* server, const dmc_volume* vol, uint64_t* ticket);
dmc_status dmc_server_poll(dmc_server* server, uint64_t ticket, int* done);
/* Block until the request is done; returns its status. */
dmc_status dmc_server_wait(dmc_server* server, uint64_t ticket, dmc_result* out);
/* As dmc_mesh_download, for a completed request. */
dmc_status dmc_server_download(dmc_server* server, uint64_t ticket, double* positions, double* normals,
                               int32_t* quads, int32_t* tris, int32_t* halfedges, int32_t* he_v, int32_t* he_f,
                               int32_t* bnd);
/* Free the request and its share of the result memory. */
dmc_status dmc_server_release(dmc_server* server, uint64_t ticket);
/* held_bytes: result blocks holding unreleased results; pool_bytes: all
 * result blocks (held and idle) plus the worker contexts' buffer pools. */
dmc_status dmc_server_stats(dmc_server* server, uint64_t* batches, uint64_t* served, size_t* held_bytes,
                            size_t* pool_bytes);
const char* dmc_server_last_error(dmc_server* server);

/* FP32 export as the CUDA reference returns it (dmc/DualMarchingCubes.cu:
 * 1076-1090): positions and normals as float triples; when negate_normals
 * is non-zero the normals are negated like the CUDA path does. */
dmc_status dmc_mesh_download_f32(dmc_context* ctx, float* positions, float* normals, int negate_normals);

/* ---- diagnostics of the last extraction's mesh (a single volume:
 * DMC_ERR_STATE after a batch); each call computes on the device and copies
 * the result out ---- */

/* Unique undirected edges, one per vertex pair
 * (dmc/HalfedgeMesh.cu:430-450, HalfedgeMesh::edges): *count receives their
 * number; edges (2 int32 per edge, may be NULL) receives each as the
 * (origin, destination) of the pair's smallest halfedge, in ascending order
 * of that halfedge.  The reference's order and orientation are those of its
 * hash table; the set of pairs is the same. */
dmc_status dmc_mesh_edges(dmc_context* ctx, int32_t* edges, int64_t* count);

/* Vertex valence distribution (dmc/VertexValence.cu:79-170,
 * VertexValence::vertexValence for quadrilaterals, or for the triangles when
 * triangles != 0, extraction with DMC_TRIANGLES): hist[k] = number of
 * vertices with k distinct incident edges (a loop counting twice), for
 * k < 100 (hist has room for 100); *len = 1 + the largest k > 0 with
 * hist[k] > 0, or 1 (the reference's trimmed size); *beyond (may be NULL) =
 * vertices of valence 100 or more, for which the reference has no bin. */
dmc_status dmc_valence_histogram(dmc_context* ctx, int triangles, uint32_t* hist, int32_t* len, int64_t* beyond);

/* Element quality (dmc/EstimateElementQuality.cu, dmc/ElementQuality.h:65-91)
 * in FP32 from the positions rounded to float.  which = 0: one value per
 * quadrilateral, the least corner ratio 2|e1 x e2| / (|e1|^2 + |e2|^2);
 * which = 1: one per triangle (extraction with DMC_TRIANGLES),
 * 2 sqrt(3) |e1 x e2| / (l0 + l1 + l2) over the squared edge lengths, 0 when
 * an edge has length 0. */
dmc_status dmc_element_quality(dmc_context* ctx, int which, float* quality);

/* The mesh as Wavefront OBJ text in cpp_dmc's main program layout
 * (cpp_dmc/main.cpp:35-87): a comment line, "v x y z" and "vn x y z" for
 * every vertex in the C++ stream's default format (printf "%g"), then
 * "f a//a b//b c//c" with 1-based indices -- triangles when triangles != 0
 * (extraction with DMC_TRIANGLES), else quadrilaterals. */
dmc_status dmc_write_obj(dmc_context* ctx, const char* path, int triangles);

/* ------------------------------------------------------------------------
 * Topologically correct Marching Cubes (the reference's other extraction
 * path).  tmc_extract replaces tmc::MarchingCubes::t_mc
 * (mc/MarchingCubes.cpp:150-318, reached from operator() at :11-147 after
 * UniformGrid::gradient) and cuda_tmc's shared-vertex extraction
 * (cuda_tmc/MarchingCubes.h:71, mc_sharedvertex; t_slice at
 * cuda_tmc/MarchingCubes.cu:1366).  Vertices on grid edges are shared and
 * numbered in the order t_mc first meets them; ambiguous cells get the
 * asymptotic-decider slicing of p_slice (mc/MarchingCubes.cpp:322-1303) with
 * its inner vertices.  Positions and normals are FP64 (normals: the
 * reference's central-difference gradients, one-sided at the border,
 * interpolated and normalised); triangles are three vertex ids each, in the
 * reference's order.  The optional orientation pass (connectivity(),
 * operator()'s c_flag) is not part of it.  flags: DMC_FIELD_ON_DEVICE and
 * DMC_STAGE_TIMING only.  The mesh stays on the device (tmc_mesh pointers)
 * until the next extraction on the context.
 * ---------------------------------------------------------------------- */
typedef struct {
    int64_t nv, nt;
    const double* d_positions;   /* 3 per vertex */
    const double* d_normals;     /* 3 per vertex */
    const int32_t* d_tris;       /* 3 vertex ids per triangle */
    int64_t active_cells;        /* cells with a cut edge */
    int64_t ambiguous_cells;     /* cells sliced with the asymptotic decider */
    int64_t face_failures;       /* of those, cells left without triangles: a face
                                    whose decider equals the iso-value and whose
                                    intersections admit no triangulation */
} tmc_mesh;

dmc_status tmc_extract(dmc_context* ctx, const dmc_grid* grid, const void* field, int field_type, double iso,
                       uint32_t flags, tmc_mesh* out);
/* Copy the last topologically correct mesh to host arrays (any may be NULL):
 * positions/normals 3*nv doubles, triangles 3*nt ints. */
dmc_status tmc_download(dmc_context* ctx, double* positions, double* normals, int32_t* triangles);

/* ------------------------------------------------------------------------
 * Launch shape and its autotuner (no counterpart in the reference, whose
 * kernels launch with fixed block sizes over one thread per item,
 * dmc/DualMarchingCubes.cu).  Results never depend on the launch shape,
 * only the time does.
 *   grid_cap     blocks (256 threads) of every grid-stride kernel, which
 *                loops over its items: 64 .. 2^24, default 32768 (env
 *                TMC_GRID_CAP);
 *   sweep_waves  waves the field sweep aims for, one 256-cell x 16-row tile
 *                of K planes each; sets K (never below 8): 256 .. 2^24,
 *                default 32768 (env TMC_SWEEP_WAVES).
 * There is no hash-table load factor to choose: vertex and quadrilateral
 * ids come from prefix sums over the cells, not from hash tables.
 * ---------------------------------------------------------------------- */
typedef struct {
    int32_t grid_cap;
    int32_t sweep_waves;
} dmc_launch;

dmc_status dmc_get_launch(dmc_context* ctx, dmc_launch* out);
/* DMC_ERR_INVALID_ARG for a value outside its range. */
dmc_status dmc_set_launch(dmc_context* ctx, const dmc_launch* launch);

typedef struct {
    dmc_launch launch;
    double ms;          /* median wall time of one dmc_extract_batch */
} dmc_tune_trial;

/* Time dmc_extract_batch(vols, nvol, field_type, flags) on ctx under
 * candidate launch shapes and keep the fastest on ctx: sweep_waves over
 * {8192, 16384, 32768, 65536} at the current grid_cap, then grid_cap over
 * {4096, 8192, 16384, 32768, 65536} at the best sweep_waves.  Each
 * candidate runs once untimed, then `reps` (1..64) times; its figure is the
 * median.  trials (optional) receives every candidate, *ntrials is its
 * capacity in and the number written out; *best (optional) the chosen
 * shape.  On an extraction error the previous shape is restored.  The
 * context's last mesh is that of a tuning run (identical under every
 * shape). */
dmc_status dmc_autotune(dmc_context* ctx, const dmc_volume* vols, int32_t nvol, int field_type, uint32_t flags,
                        int32_t reps, dmc_tune_trial* trials, int32_t* ntrials, dmc_tune_trial* best);

/* Per-stage device times (ms) of the last extraction run with
 * DMC_STAGE_TIMING.  names[i] point to static strings.  *count is in/out:
 * capacity in, number of stages out. */
dmc_status dmc_stage_times(dmc_context* ctx, const char** names, float* ms, int* count);

/* Device memory helpers for callers without their own HIP runtime. */
dmc_status dmc_device_alloc(dmc_context* ctx, size_t bytes, void** ptr);
dmc_status dmc_device_free(dmc_context* ctx, void* ptr);
dmc_status dmc_copy_to_device(dmc_context* ctx, void* dst, const void* src, size_t bytes);
dmc_status dmc_copy_to_host(dmc_context* ctx, void* dst, const void* src, size_t bytes);
dmc_status dmc_synchronize(dmc_context* ctx);

const char* dmc_status_str(dmc_status s);
/* Text of the last error on this context (empty string if none). */
const char* dmc_last_error(dmc_context* ctx);

#ifdef __cplusplus
}
#endif
#endif
