/*
 * horayzon.h -- C-ABI of libhorayzon.so, the MI355X (gfx950) horizon and
 * sky-view-factor library.
 *
 * Plain C types only (pointers, ints, doubles): a Python ctypes stub, a
 * Cython extern block or a Fortran ISO_C_BINDING interface binds it
 * directly.  All entry points but hz_walk_noalloc are synchronous: they
 * return once the results are in the caller's arrays (or in device memory,
 * for hz_domain_run).
 *
 * Error convention: 0 on success, a negative HZ_E* code on failure;
 * hz_last_error() returns the message of the calling thread's last failure.
 * On failure the caller's output arrays are filled with NaN, never left
 * half-written.
 */
#ifndef HORAYZON_H
#define HORAYZON_H

#ifdef __cplusplus
extern "C" {
#endif

#define HZ_OK 0
#define HZ_EINVAL (-1)  /* argument out of the reference's documented range */
#define HZ_EDEVICE (-2) /* HIP runtime error (no device, launch failure, ...) */
#define HZ_ENOMEM (-3)  /* device or host allocation failed */
#define HZ_EKERNEL (-4) /* a kernel reported an internal limit (e.g. stack depth) */

/* Search strategies of the elevation-angle search. */
#define HZ_PROLOGUE_HOST 0
#define HZ_PROLOGUE_DEVICE 1

#define HZ_MODE_WALK 0    /* the reference walk: carried elevation, 0.5 deg steps, -85 deg floor */
#define HZ_MODE_CLIMB 1   /* clip-aware per-azimuth climb from 0 deg (independent azimuths) */
#define HZ_MODE_LEARNED 2 /* lattice search started at a predicted horizon (needs a predictor
                             and num_hori * refine_factor == 24) */

typedef struct hz_options {
    int search_mode;      /* HZ_MODE_*; default HZ_MODE_WALK */
    int leaf_size;        /* BVH leaf size 1..8; 0 = library default (1: every leaf is one
                             triangle stored inline in its parent node) */
    int block_size;       /* trace kernel workgroup size (64, 128, 256); 0 = default */
    int collect_counters; /* 1: count BVH node visits and triangle tests (slower) */
    int cell_order;       /* 1: trace cells in Morton order of their position (default), 0: input order */
    int verbose;          /* 1: print the reference's progress lines to stderr */
    int trace_kernel;     /* 0: persistent lanes with per-XCD work queues (default),
                             1: one lane per cell with azimuths in lock step */
    int refill;           /* persistent kernel: refill lanes once fewer than this many of
                             the 64 lanes of a wavefront hold a live query (1..64; 0 = default, 36) */
    int bvh_morton;       /* LBVH codes: 0 = 3D in a cube of the largest extent, 1 = 2D x/y */
    int queue_chunk;      /* persistent kernel: most cells a wavefront takes per work-queue
                             fetch (it takes as many as its lanes ask for; 0 = 64) */
    int bvh_width;        /* traversal tree: 2 = binary nodes with full-precision boxes
                             (default), 4 = four-wide nodes with 8-bit quantized child boxes
                             (half the node fetches, same results); the lane-per-cell kernel
                             always uses 2 */
    int prologue;         /* where the FP64 mesh / ENU / frame prologue runs:
                             HZ_PROLOGUE_HOST (default): host threads, glibc sin / cos, the
                             reference's values bit for bit;
                             HZ_PROLOGUE_DEVICE: on the GPU from the raw ICON arrays, with
                             correctly rounded sin / cos (glibc's differ in the last bit on
                             ~0.2% of arguments); the ENU origin is still the host's */
    int whole_cells;      /* lattice searches (HZ_MODE_CLIMB, HZ_MODE_LEARNED) in the
                             persistent kernel: 0 (default) = one work item per (cell,
                             azimuth), 1 = one per cell (as the walk); same results */
    int certify;          /* lattice searches (HZ_MODE_CLIMB, HZ_MODE_LEARNED): 1 (default) =
                             equal to the reference walk: cells with a mesh-boundary point
                             (domain border, holes at vertices with fewer than six cells)
                             within reach run the reference walk itself, bit for bit; every
                             other cell sees hits monotone in elevation, where the lattice
                             search returns the walk's values (to the walk's FP64 rounding of
                             its elevation, ~1e-14 deg); 0 = the plain lattice search, which
                             differs from the walk where its rays pass through mesh holes.
                             Needs (cell, azimuth) work items (persistent kernel, whole_cells
                             0); without memory for them the walk runs everywhere */
    int cache;            /* one-shot calls (horizon_svf_comp_ex with the host prologue): 1
                             (default) = keep the device buffers, streams and pinned staging
                             of the call for the next call on the device whose grid has the
                             same sizes (cells, triangles, mesh vertices), which then only
                             uploads its scene; a call with other sizes frees them first.
                             0 = free everything at the end of the call.  hz_cache_release
                             frees what is kept.  The entry points without options use 1 */
    int bvh_sah;          /* 1: rebuild the top of the LBVH (the nodes over more than
                             max(leaf_size, 8, triangles / 8192) triangles) with a binned
                             surface-area heuristic over the subtrees below it (binary trees
                             of 4096 or more triangles; the build then waits for the host
                             once); 0 (default): the Morton tree.  Same results either way */
} hz_options;

typedef struct hz_stats {
    long long num_rays;          /* occlusion queries cast */
    long long node_visits;       /* BVH interior nodes fetched (collect_counters only) */
    long long tri_tests;         /* ray/triangle tests (collect_counters only) */
    long long num_triangles;     /* triangles in the mesh */
    long long num_mesh_vertices; /* mesh vertices (cells + added ICON vertices) */
    long long bvh_nodes;         /* interior BVH nodes */
    int bvh_depth;               /* deepest interior level (0 = root) */
    int kernel_error;            /* kernel-side error flags (0 = none) */
    double t_prologue_ms;        /* host FP64 mesh + ENU + frames */
    double t_upload_ms;          /* host -> device copies */
    double t_bvh_ms;             /* device LBVH build (HIP events) */
    double t_order_ms;           /* device cell ordering (HIP events) */
    double t_trace_ms;           /* device search + fused sector mean / SVF (HIP events) */
    double t_download_ms;        /* device -> host copies */
    double t_total_ms;           /* wall time of the call */
    double lon0, lat0;           /* ENU origin [rad] */
    double t_predict_ms;         /* device raster + descriptors + predictor (learned mode) */
    double refill_frac;          /* collect_counters, persistent kernel: share of wavefront
                                    time spent refilling lanes (search state, next query) */
    double lanes_per_iter;       /* collect_counters, persistent kernel: live lanes (of 64) per
                                    traversal iteration */
    double cert_walk_cells;      /* certified lattice searches: cells that ran the reference
                                    walk (a mesh-boundary point in reach); the others ran the
                                    lattice search */
    double t_certify_ms;         /* certified lattice searches: classification of the cells
                                    (before the BVH build; HIP events) */
    double t_cert_walk_ms;       /* certified lattice searches: the walk of the near cells */
    double t_border_walk_ms;     /* certified lattice searches: the team walk of the cells next
                                    to mesh holes (beside the other two launches) */
    double spec_rays;            /* queries the team walk cast ahead and did not use (not in
                                    num_rays) */
    double t_cert_lattice_ms;    /* certified lattice searches: the far cells' lattice search */
} hz_stats;

/*
 * Drop-in for the reference entry point
 *   void horizon_svf_comp(...)   mo_lradtopo_horayzon.h:4-11,
 *                                mo_lradtopo_horayzon.cpp:562-570
 * Same argument order, types, units and output layout:
 *   horizon[j * num_cell + i]  (deg), skyview[i]  (-).
 * Errors are printed to stderr and the outputs are filled with NaN.
 */
void horizon_svf_comp(double* clon, double* clat, double* hsurf, double* vlon, double* vlat,
                      int* cells_of_vertex, double* horizon, double* skyview, int num_cell,
                      int num_vertex, int num_hori, int grid_type, double dist_search,
                      double ray_org_elev, int refine_factor, int svf_type);

/*
 * horizon_svf_comp with options, a status code and statistics; the
 * counterpart of the same reference call (mo_lradtopo_horayzon.cpp:562-785),
 * whose only reporting is the printed timers and ray counts (:604-608,
 * :424-428, :763-773).  opt and stats may be NULL.
 */
int horizon_svf_comp_ex(const double* clon, const double* clat, const double* hsurf,
                        const double* vlon, const double* vlat, const int* cells_of_vertex,
                        double* horizon, double* skyview, int num_cell, int num_vertex,
                        int num_hori, int grid_type, double dist_search, double ray_org_elev,
                        int refine_factor, int svf_type, const hz_options* opt,
                        hz_stats* stats);

/*
 * Fortran-interface variant of horizon_svf_comp: the reference's commented
 * EXTPAR interface (mo_lradtopo_horayzon.cpp:561-574, :778-786) adds
 * `char* buffer, int* buffer_len` and copies what it printed to stdout into
 * the caller's buffer.  Here the same report (mesh, ENU origin, SVF form, BVH
 * and tracing times, rays per cell and sector, or the failure) is written to
 * buffer: at most *buffer_len bytes, not NUL-terminated (a Fortran
 * CHARACTER), and *buffer_len is set to the number of bytes written.
 * Outputs are NaN on failure, as for horizon_svf_comp.
 */
void horizon_svf_comp_fortran(double* clon, double* clat, double* hsurf, double* vlon,
                              double* vlat, int* cells_of_vertex, double* horizon,
                              double* skyview, int num_cell, int num_vertex, int num_hori,
                              int grid_type, double dist_search, double ray_org_elev,
                              int refine_factor, int svf_type, char* buffer, int* buffer_len);

/*
 * Frees the device buffers the one-shot calls keep for the next call of the
 * same grid size (option cache).  No counterpart in the reference, which
 * builds and frees its scene in every call (mo_lradtopo_horayzon.cpp:395-432,
 * :780-785).
 */
void hz_cache_release(void);

/* Message of the calling thread's last failure ("" if none). */
const char* hz_last_error(void);

/* Fill *opt with the library defaults. */
void hz_default_options(hz_options* opt);

/*
 * Resident domain: the FP64 prologue (mesh, ENU, per-cell frames) runs once
 * in hz_domain_create and its FP32 scene and frames stay in device memory;
 * hz_domain_run then performs one full device pass (LBVH build, cell
 * ordering, search, sector mean and SVF) with its outputs kept in device
 * memory until hz_domain_fetch copies them out.  The three calls together
 * compute exactly what horizon_svf_comp computes (reference
 * mo_lradtopo_horayzon.cpp:596-761); the split exists so that callers can
 * time the device pass alone or re-run it with other search settings.
 */
typedef struct hz_domain hz_domain;

int hz_domain_create(const double* clon, const double* clat, const double* hsurf,
                     const double* vlon, const double* vlat, const int* cells_of_vertex,
                     int num_cell, int num_vertex, int grid_type, double ray_org_elev,
                     const hz_options* opt, hz_domain** out, hz_stats* stats);

int hz_domain_run(hz_domain* dom, int num_hori, double dist_search, int refine_factor,
                  int svf_type, const hz_options* opt, hz_stats* stats);

int hz_domain_fetch(hz_domain* dom, double* horizon, double* skyview, hz_stats* stats);

void hz_domain_destroy(hz_domain* dom);

/*
 * Diagnostics, host only (no device needed): the FP32 ENU scene and per-cell
 * ray frames that the device pass traces, i.e. the output of the reference's
 * mesh builder and FP64 prologue (mo_lradtopo_horayzon.cpp:596-722; the
 * reference's own check of this stage is the printed mesh checksum of
 * verification.py:171-172).  Pass NULL arrays to obtain the counts only.
 *   vert32[3 * num_mesh_vertices], tri[3 * num_triangles],
 *   frame_n / frame_north [3 * num_cell] (FP64), org [3 * num_cell] (FP32),
 *   lon0_lat0[2] (ENU origin, rad)
 */
int hz_scene_export(const double* clon, const double* clat, const double* hsurf,
                    const double* vlon, const double* vlat, const int* cells_of_vertex,
                    int num_cell, int num_vertex, int grid_type, double ray_org_elev,
                    long long* num_triangles, long long* num_mesh_vertices, float* vert32,
                    int* tri, double* frame_n, double* frame_north, float* org,
                    double* lon0_lat0);

/*
 * Terrain slope and aspect [deg] per cell (no counterpart in the reference,
 * whose sky view factor is for a horizontal plane,
 * mo_lradtopo_horayzon.cpp:320-332): the surface normal is the area-weighted
 * normal of the mesh triangles around the cell's circumcentre; slope is its
 * angle to the local vertical, aspect the azimuth (clockwise from north,
 * [0, 360)) of its horizontal part, i.e. the downhill direction; 0 / 0 for a
 * level cell or a cell that is no triangle corner.
 */
int hz_domain_slope_aspect(hz_domain* dom, double* slope, double* aspect);

/*
 * Learned elevation bracket (no counterpart in the reference, which always
 * runs the full search, mo_lradtopo_horayzon.cpp:490-555).  A predictor is
 * a 192 -> 256 -> 256 -> 24 GELU network evaluated in BF16 on the matrix
 * cores; its input is the per-cell descriptor (24 azimuths x 8 radial rings
 * of maximum terrain elevation angle, 0.5 .. 40 km), its output the horizon
 * per azimuth [deg].  HZ_MODE_LEARNED starts each azimuth's lattice search at
 * the predicted elevation: it climbs while the query hits and descends while
 * it misses, so a wrong prediction costs queries, never accuracy; with the
 * default option certify = 1 its results equal the reference walk's.
 *
 * params layout (hz_predictor_num_params() floats): w1[256][192], b1[256],
 * w2[256][256], b2[256], w3[24][256], b3[24].
 */
typedef struct hz_predictor hz_predictor;
long long hz_predictor_num_params(void);
int hz_predictor_create(const float* params, hz_predictor** out);
void hz_predictor_destroy(hz_predictor* p);
/* The predictor must outlive the runs that use it. */
int hz_domain_set_predictor(hz_domain* dom, const hz_predictor* p);
/* Descriptors of all cells, BF16 bit patterns, desc[cell * 192 + az * 8 + ring]. */
int hz_domain_descriptors(hz_domain* dom, unsigned short* desc);
/* Predictions of the last learned run, pred[cell * 24 + az] [deg]. */
int hz_domain_predictions(hz_domain* dom, float* pred);

/*
 * Native trainer of the predictor (Adam, mean squared error, cosine step
 * decay), run on the device with the same MFMA kernels.  X: BF16 bit
 * patterns [n][192]; T: targets [n][24] in deg.
 */
typedef struct hz_trainer hz_trainer;
int hz_trainer_create(unsigned long long seed, hz_trainer** out);
void hz_trainer_destroy(hz_trainer* t);
int hz_trainer_set_weights(hz_trainer* t, const float* params);
int hz_trainer_get_weights(hz_trainer* t, float* params);
int hz_trainer_set_data(hz_trainer* t, const unsigned short* X, const float* T, int n);
/* epoch_loss (optional, epochs floats): loss of the last batch of each epoch */
int hz_trainer_fit(hz_trainer* t, int epochs, int batch, float lr, float* epoch_loss);
/* One forward + backward pass on (X, T) without an update: gradients of the
 * loss in params layout, and the loss.  n must be a multiple of 8. */
int hz_trainer_gradients(hz_trainer* t, const unsigned short* X, const float* T, int n,
                         float* grads, float* loss);
int hz_trainer_predict(hz_trainer* t, const unsigned short* X, int n, float* out);

/*
 * Batched domains (no counterpart in the reference, which handles one domain
 * per call): one persistent trace launch over the cells of many resident
 * domains that share the search parameters.  Per run, every domain's LBVH
 * and cell order are rebuilt (stream-ordered), then the trace launch runs --
 * replayed from a captured HIP graph when use_graph is set (the graph is
 * captured on the first run of a configuration).  Results stay in each
 * domain; hz_domain_fetch reads them.  Domains must outlive the batch.
 * Options: search_mode (walk or climb), leaf_size, block_size, refill,
 * cell_order, bvh_morton, bvh_width and queue_chunk.
 */
typedef struct hz_batch hz_batch;
int hz_batch_create(hz_domain* const* domains, int n, hz_batch** out);
int hz_batch_run(hz_batch* b, int num_hori, double dist_search, int refine_factor, int svf_type,
                 const hz_options* opt, int use_graph, hz_stats* stats);
void hz_batch_destroy(hz_batch* b);

/* Number of cells / mesh triangles of a resident domain. */
int hz_domain_num_cell(const hz_domain* dom);
long long hz_domain_num_triangles(const hz_domain* dom);

/*
 * The scene a resident domain traces, read back from the device (either
 * prologue): vert32[3 * num_mesh_vertices], tri[3 * num_triangles],
 * frame_n / frame_north [3 * num_cell] (FP64), org [3 * num_cell] (FP32).
 * NULL arrays are skipped.  Same layout as hz_scene_export, which computes
 * it on the host; the two are compared to check the device prologue.  No
 * reference counterpart (the reference keeps its scene inside Embree, :406-414).
 */
int hz_domain_scene(const hz_domain* dom, float* vert32, int* tri, double* frame_n,
                    double* frame_north, float* org);

/*
 * Builds the traversal tree of a resident domain with the tree options of
 * opt (leaf_size, bvh_width, bvh_morton; NULL: defaults) and reads it back:
 * nodes[128 * num_nodes] bytes of BvhNode (width 2) or Bvh4Node (width 4)
 * records, root first, and tris[12 * num_tris] floats, the triangles in
 * leaf order (v0 xyz, v1 xyz, v2 xyz, original index as int bits, 2 pad).
 * NULL arrays are skipped (call once for the counts).  A binary node's z
 * planes bound z' = z + (x^2 + y^2) / (2 * 6371229), the height above the
 * sphere touching the ENU plane at its origin, and a node whose children are
 * both interior holds their leaf-order split as int bits in tri[1].  For
 * checking the tree's structure; the reference's counterpart is Embree's internal build
 * (rtcCommitScene, mo_lradtopo_horayzon.cpp:422), which it cannot read back.
 */
int hz_domain_bvh(hz_domain* dom, const hz_options* opt, void* nodes, long long* num_nodes,
                  float* tris, long long* num_tris);

/*
 * The device prologue's correctly rounded sin and cos (the replacement of
 * the glibc sincos calls of lonlat2ecef, mo_lradtopo_horayzon.cpp:158-169),
 * for n arguments with |x| < 2^20: on the device (on_device = 1) or the same
 * code on the host (0).  For checking the routine against exact values.
 */
int hz_sincos(const double* x, long long n, double* s, double* c, int on_device);

/*
 * Measurement support, no counterpart in the reference: the rate, in steps
 * per second, of dependent random reads of 64 bytes (a traversal step's node
 * boxes) from num_nodes 128-byte records, at the trace kernel's occupancy
 * and lane activity; hot_nodes of the records take every other step.  The
 * trace kernel's node fetches per second against this rate place it on the
 * GPU's random-gather roofline (bench.py).  Allocates num_nodes * 128 bytes
 * of device memory for the call; num_nodes in [1024, 2^27], iters (steps
 * per lane) in [1, 100000].
 */
int hz_calibrate_gather(long long num_nodes, long long hot_nodes, int iters, double* rate);

/*
 * The reference walk without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * It computes what hz_domain_run computes with search_mode 0 (the result of
 * horizon_svf_comp, bit for bit) for a scene the caller already holds in
 * device memory, in the layout hz_scene_export writes:
 *   vert32[3 * num_mesh_vertices], tri[3 * num_triangles] (indices into
 *   vert32, which the call does not check), frame_n / frame_north
 *   [3 * num_cell] (FP64), org[3 * num_cell] (FP32),
 * into the caller's device outputs horizon[num_hori * num_cell] (deg, row j =
 * azimuth j) and skyview[num_cell].  Every intermediate buffer lives in one
 * static device array of HZ_WALK_NOALLOC_SCRATCH_MIB MiB, which each device
 * the call runs on holds until the process ends; scenes of more than
 * HZ_WALK_NOALLOC_MAX_TRIANGLES triangles or HZ_WALK_NOALLOC_MAX_CELLS cells
 * are refused with HZ_ENOMEM.  The work is enqueued on stream (a
 * hipStream_t; NULL: the null stream) and the call returns without waiting:
 * synchronize the stream before reading the outputs.  The first call on a
 * device also writes a constant table into the array with a synchronous
 * copy, so make it outside stream capture.  All calls on one device share
 * the array: calls that overlap on one device race.
 */
#define HZ_WALK_NOALLOC_MAX_TRIANGLES 196608
#define HZ_WALK_NOALLOC_MAX_CELLS 65536
#define HZ_WALK_NOALLOC_SCRATCH_MIB 96
int hz_walk_noalloc(const float* vert32, long long num_mesh_vertices, const int* tri,
                    long long num_triangles, const double* frame_n, const double* frame_north,
                    const float* org, int num_cell, int num_hori, double dist_search,
                    int refine_factor, int svf_type, double* horizon, double* skyview,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HORAYZON_H */
