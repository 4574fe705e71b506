// Internal declarations shared by the host prologue, the LBVH builder, the
// trace kernels and the C-ABI layer of libhorayzon.so.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <string>
#include <memory>
#include <utility>
#include <vector>

namespace hz {

// ---------------------------------------------------------------------------
// Device-side records
// ---------------------------------------------------------------------------

// Per-cell ray frame: unit sphere normal n and unit north direction (FP64,
// reference mo_lradtopo_horayzon.cpp:693-712) and the FP32 ray origin
// (:715-720).  64 bytes, one record per cell.  flags bit 0: the cell is a
// corner of an ICON vertex with fewer than six cells (next to a mesh hole).
struct alignas(16) CellFrame {
    double n[3];
    double north[3];
    float org[3];
    int pad;  // flags
};
static_assert(sizeof(CellFrame) == 64, "CellFrame layout");

// BVH2 interior node, 128 bytes (one L2 line): the conservative boxes of
// both children, the two child references and, for a child that is a
// single-triangle leaf, that triangle's FP32 vertices inline, so one fetch
// brings everything a step needs.
//   ref >= 0 : interior node index in bits 0..27; bit 30 set when that node
//              holds an inline triangle (its second 64 bytes must be read)
//   ref <  0 : leaf; ~ref = first | ((count - 1) << 28), triangles in leaf
//              order; a leaf of count 1 is the inline triangle of its slot
// The z planes of a binary tree's boxes bound z' = z + (x^2 + y^2) * kZCurv,
// the height above the sphere of radius 1 / (2 kZCurv) that touches the ENU
// plane at its origin (to second order), not z: terrain far from the origin
// drops below the ENU plane and tilts against it, and its boxes in z' stay
// flat (the traversal tests a ray's z' range over its x/y interval).
// A node whose children are both interior has no inline triangle; its
// tri[1] then holds (as int bits) the leaf-order index of child 0's last
// leaf, child 1's leaves following it, or -1 where the children's leaves are
// not two such ranges (the top rebuilt with the SAH option).
constexpr double kZCurv = 1.0 / (2.0 * 6371229.0);  // Earth radius, reference :584
struct alignas(16) BvhNode {
    float lx0, lx1, hx0, hx1;
    float ly0, ly1, hy0, hy1;
    float lz0, lz1, hz0, hz1;
    int r0, r1;
    float tri[18];  // slot 0: tri[0..8], slot 1: tri[9..17] (v0, v1, v2 xyz)
};
static_assert(sizeof(BvhNode) == 128, "BvhNode layout");

// Four-wide node, 128 bytes, built by collapsing the binary LBVH (every
// binary interior node at even depth becomes one node whose children are its
// grandchildren, or its children where those are leaves).  Child boxes are
// quantized to 8 bits per plane relative to the node's own box and rounded
// outward: plane = org + q * 2^(e - 127).
//   org[3], meta   : FP32 origin; meta bytes 0..2 = biased exponents e of
//                    x, y, z, byte 3 = number of children (1..4, packed first)
//   q[6]           : lo x, hi x, lo y, hi y, lo z, hi z; byte k = child k
//   ref[4]         : >= 0 interior node (bits 0..27), bit 30 set when that
//                    node holds inline triangles; < 0 leaf, ~ref = code:
//                    bit 30 set: inline triangle, slot = bit 0, leaf-order
//                    index in bits 1..27;
//                    else first | ((count - 1) << 27), triangles in leaf order
//   tri[18]        : inline slots 0 (tri[0..8]) and 1 (tri[9..17])
// The first 64 bytes hold everything but the inline triangles (and the first
// two floats of slot 0), so a node without them is read as 64 bytes.
struct alignas(16) Bvh4Node {
    float org[3];
    unsigned meta;
    unsigned q[6];
    int ref[4];
    float tri[18];
};
static_assert(sizeof(Bvh4Node) == 128, "Bvh4Node layout");
constexpr int kB4InlineBit = 1 << 30;
constexpr int kB4CountShift = 27;
constexpr int kB4FirstMask = (1 << kB4CountShift) - 1;
constexpr int kMaxTriangles4 = 1 << kB4CountShift;

constexpr int kLeafFirstBits = 28;
constexpr int kLeafFirstMask = (1 << kLeafFirstBits) - 1;
constexpr int kWideBit = 1 << 30;
constexpr int kMaxLeafSize = 8;

// Triangles in leaf order (used by leaves of more than one triangle):
// 3 float4 per triangle {v0.x, v0.y, v0.z, v1.x}, {v1.y, v1.z, v2.x, v2.y},
// {v2.z, id, 0, 0}
struct DeviceBvh {
    BvhNode* nodes = nullptr;  // Bvh4Node records when width == 4 (same 128-byte stride)
    float4* tris = nullptr;
    int num_tri = 0;
    int num_nodes = 0;  // interior nodes (>= 1 whenever num_tri >= 1)
    int depth = 0;      // deepest interior level, root = 0
    int leaf_size = 1;  // largest leaf (> 1: leaves beyond the inline triangles exist)
    int width = 2;      // 2: BvhNode, 4: Bvh4Node
    const unsigned* d_num_nodes = nullptr;  // device count of four-wide nodes
};

// Constants of the elevation search (reference :577-587, :651-668).
struct WalkParams {
    double hori_acc;        // 0.25 deg [rad]
    double elev_thresh;     // -85 deg [rad]
    double azim_sin, azim_cos;  // of 360 deg / (nhori * rf)
    double elev_sin_2ha, elev_cos_2ha;
    double shift_sin, shift_cos;  // of -azim_shift
    float tfar;             // (float) dist_search
    float tfar_box;         // slab-test far limit: tfar * 1.0001 + 1 m
    int num_hori;
    int refine_factor;
    int svf_type;
    int ncell;
    // lattice search (climb / learned modes): cos, sin of k * 0.5 deg for
    // k = 0 .. kMaxLattice, and the predicted horizon [deg] per cell
    // (pred[cell * pred_stride + azimuth]) of the learned mode
    const double2* elev_cs;
    const float* pred;
    int pred_stride;
    // sin((2k + 1) * hori_acc), k = 0 .. kMaxLattice: every horizon value is
    // (2k + 1) * hori_acc (to the walk's accumulated rounding) or 0
    const double* sin_h;
    // launch only: workgroups the persistent item launch leaves free for a
    // launch running beside it (the walk of the cells next to mesh holes)
    int reserve_blocks;
    // device {first, count}: the item kernels, k_items_finish and k_certify
    // work on these positions of the trace order only (NULL: all cells)
    const int* range;
    // binary trees: per cell, the reference of the node its queries start
    // at (launch_start_nodes; NULL: the root)
    const int* start;
};

constexpr int kMaxLattice = 180;  // 90 deg

// ---------------------------------------------------------------------------
// Where the lattice searches equal the reference walk (certify.hip): cells
// with no mesh-boundary point within reach of their rays see hits that are
// monotone in elevation, so the lattice search returns the walk's values
// there; the certified searches run the walk on every other cell.
// ---------------------------------------------------------------------------

// Mesh-boundary points of a domain in a uniform x/y grid (CSR).  Built once
// per domain and search distance.
struct CertIndex {
    float4* pts = nullptr;           // boundary points (FP32 ENU x, y, z), bucket order
    int* off = nullptr;              // nx * ny + 1 bucket offsets
    unsigned char* near = nullptr;   // per bucket: a boundary point may be in reach
    int npts = 0, nx = 0, ny = 0, rb = 0;  // rb: bucket radius scanned around a cell
    double x0 = 0.0, y0 = 0.0, S = 1.0;    // grid origin and bucket size [m]
    double lw = 0.0;                 // longest boundary edge (+ margin) [m]
    double reach = 0.0;              // (double)(float)dist_search [m]
    double r2max = 0.0;              // squared horizontal distance of a boundary point in
                                     // reach: (reach + lw)^2 + lw^2
    float dist_key = -1.0f;          // dist_search the index was built for
};

// Boundary points (cells next to a vertex with fewer than six cells, i.e.
// CellFrame flag bit 0) and the longest triangle edge between two of them:
// gathered on the device, bucketed on the host.  hmin / hmax: range of the
// cell elevations (bounds the vertical offset of a boundary point).
int cert_index_build(CertIndex& ci, const CellFrame* cells, const float* vert, int ncell,
                     const int* tri, long long ntri, const float bounds_xy[4], double hmin,
                     double hmax, float tfar, hipStream_t st, std::string& err);
void cert_index_free(CertIndex& ci);
// Flag bit 1 of the cells with a mesh-boundary point in reach (bit 0 cells
// excluded), or clear it (clear_only); counts_ranges (10 ints) receives two
// class counts and the {first, count} position ranges of the three classes
// of the trace order that sorts by them (cells next to mesh holes, cells
// with bit 1, the rest) and of the last two together.
void launch_cert_classes(const CertIndex& ci, CellFrame* cells, int ncell, bool clear_only,
                         int* counts_ranges, hipStream_t st);

// cos / sin of k * (2 * hori_acc), k = 0 .. kMaxLattice, through glibc
// sincos, and sin((2k + 1) * hori_acc) through glibc sin (host)
void lattice_table(double hori_acc, double2* out, double* sin_h);

// ---------------------------------------------------------------------------
// Host prologue (prologue.cpp)
// ---------------------------------------------------------------------------

// Allocator that leaves trivially constructible elements uninitialised on
// resize: large host buffers are then first touched by the threads that fill
// them, instead of being zeroed by one thread first.
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U>&) {}
    template <class U>
    void construct(U* p) noexcept {
        ::new ((void*)p) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using HostVec = std::vector<T, UninitAlloc<T>>;

struct HostScene {
    HostVec<float> vert32;  // 3 floats per mesh vertex (ENU, FP32)
    HostVec<int> tri;       // 3 vertex indices per triangle
    HostVec<CellFrame> frame;
    long long num_mesh_vertices = 0;
    double lon0 = 0.0, lat0 = 0.0;
    float bounds_xy[4] = {0.f, 0.f, 0.f, 0.f};  // x / y extent of vert32 (min x, min y, max x, max y)
    HostVec<double> work;   // FP64 mesh vertices (lon/lat/h, then ECEF, then ENU)
    HostVec<int> vcount;    // per ICON vertex: triangles, added mesh vertices
    HostVec<long long> voff;  // their prefix sums
    HostVec<int> holes;       // cells next to mesh holes (hole_cells)
    double centre[3] = {0.0, 0.0, 0.0};  // earth centre and north pole in ENU
    double pole[3] = {0.0, 0.0, 0.0};
};

int prologue(const double* clon, const double* clat, const double* hsurf, const double* vlon,
             const double* vlat, const int* cells_of_vertex, int num_cell, int num_vertex,
             int grid_type, double ray_org_elev, HostScene& out, std::string& err);
// The two halves of prologue(): the mesh, ECEF and ENU transforms and the FP32
// scene (vert32, tri, bounds, ENU origin), then the per-cell frames and hole
// flags (frame), which read the FP64 ENU cell points prologue_scene leaves in
// out.work.  A caller can start the tree build between the two.
int prologue_scene(const double* clon, const double* clat, const double* hsurf,
                   const double* vlon, const double* vlat, const int* cells_of_vertex,
                   int num_cell, int num_vertex, int grid_type, HostScene& out,
                   std::string& err);
void prologue_frames(const int* cells_of_vertex, int num_cell, int num_vertex,
                     double ray_org_elev, HostScene& out);
// The cells next to mesh holes (corners of an ICON vertex with fewer than six
// cells), the cells prologue_frames flags with bit 0; a cell may repeat.
void hole_cells(const int* cells_of_vertex, int num_cell, int num_vertex, HostVec<int>& out);
// The frames of prologue_frames computed on the device from the FP64 ENU cell
// points P (3 per cell, as prologue_scene leaves them), with the same FP64
// operations in the same order (IEEE add, multiply, divide and square root:
// the same bits), flags clear; launch_flag_cells then sets flag bit 0 on
// the listed cells.  Enqueued on st.
void launch_frames_from_enu(const double* P, int ncell, const double centre[3],
                            const double pole[3], double ray_org_elev, CellFrame* cells,
                            hipStream_t st);
void launch_flag_cells(const int* list, int n, CellFrame* cells, hipStream_t st);

WalkParams make_walk_params(int num_hori, int refine_factor, int svf_type, double dist_search,
                            int ncell);

// ENU origin and transform constants (reference :178-201, :624-644): the
// origin is the centroid of the cell points summed in cell order; so/co and
// sa/ca are sin/cos of its longitude / latitude; (x0, y0, z0) the surface
// point below it; centre / pole the earth centre and north pole in ENU.
struct EnuFrame {
    double lon0, lat0;
    double so, co, sa, ca;
    double x0, y0, z0;
    double centre[3], pole[3];
};
// from the ECEF cell points P[3 * i .. 3 * i + 2], i < ncell
void enu_frame_from_sum(const double* P, int ncell, EnuFrame& E);
// from the cell lon / lat / elevation (host sin / cos, as the host prologue)
void enu_frame_host(const double* clon, const double* clat, const double* hsurf, int ncell,
                    EnuFrame& E);

// Host worker pool (pool.cpp).  pool_submit queues f as part of group grp;
// pool_wait returns once every item of grp has run, running queued items
// itself meanwhile.  Items must not wait on other items.  pool_for runs
// f(b, e) over [0, n) in up to pool_threads() ranges of at least
// min_per_task indices, the caller taking the first.
struct TaskGroup {
    std::atomic<int> left{0};
};
int pool_threads();
void pool_submit(TaskGroup& grp, std::function<void()> f);
void pool_wait(TaskGroup& grp);
void pool_for(long long n, const std::function<void(long long, long long)>& f,
              long long min_per_task = 4096);

// Large host <-> device copies through pinned slots (stage.cpp): copy()
// moves `bytes` between pageable host memory and device memory on stream
// st and returns once the copy is complete.  One Stager per stream user;
// its pinned slots are allocated on first use and kept until release().
struct Stager {
    static constexpr int kSlots = 3;
    void* buf = nullptr;
    hipEvent_t ev[kSlots] = {nullptr, nullptr, nullptr};
    hipError_t copy(void* dst, const void* src, size_t bytes, bool to_device, hipStream_t st);
    void release();
};

// Device prologue (prologue_dev.hip): the same mesh, FP32 scene and frames
// as prologue(), built on the device from the raw ICON arrays.  Longitude /
// latitude sin and cos are correctly rounded (crtrig.h); the ENU origin is
// enu_frame_host's, computed on host threads while the device works.
struct DeviceScene {
    float* vert = nullptr;      // FP32 ENU mesh vertices, 3 per vertex
    int* tri = nullptr;         // 3 mesh vertex indices per triangle
    CellFrame* cells = nullptr; // per-cell frame, origin and flags
    long long ntri = 0, nmv = 0;
    double lon0 = 0.0, lat0 = 0.0;
    float bounds_xy[4] = {0.f, 0.f, 0.f, 0.f};  // x/y extent of the vertices
};
// Arrays of `out` are allocated here (hipMalloc); the caller frees them.
int prologue_device(const double* clon, const double* clat, const double* hsurf,
                    const double* vlon, const double* vlat, const int* cells_of_vertex,
                    int num_cell, int num_vertex, int grid_type, double ray_org_elev,
                    hipStream_t st, Stager& sg, DeviceScene& out, std::string& err);
// sin / cos of n arguments with the device prologue's correctly rounded
// routine, on the device (device != 0) or on the host
int sincos_cr_eval(const double* x, long long n, double* s, double* c, int device,
                   std::string& err);

// Cell -> adjacent triangles (CSR, triangles in increasing order): the
// triangles with cell c (mesh vertex c < ncell) as a corner are
// adj[off[c] .. off[c + 1]).  Needed by slope / aspect only.
void cell_adjacency(const int* tri, long long ntri, int ncell, std::vector<int>& off,
                    std::vector<int>& adj);

// Slope / aspect [deg] per cell (terrain.hip)
void slope_aspect(const float* vert, const int* tri, const int* off, const int* adj,
                  const CellFrame* cells, int ncell, double* slope, double* aspect,
                  hipStream_t st);

// ---------------------------------------------------------------------------
// LBVH builder (lbvh.hip)
// ---------------------------------------------------------------------------

struct LbvhWork {
    int capacity = 0;  // triangles the buffers hold
    uint64_t* keys[2] = {nullptr, nullptr};
    int* vals[2] = {nullptr, nullptr};
    float4* aabb_lo = nullptr;  // leaves [0, n), interior [n, 2n-1)
    float4* aabb_hi = nullptr;
    int* child = nullptr;   // 2 per interior node
    int* range = nullptr;   // 2 per interior node (first, last leaf)
    int* parent = nullptr;  // leaves [0, n), interior [n, 2n-1)
    void* boxes = nullptr;  // range-union work: sorted leaf boxes, prefix, suffix, sparse table
    float* part = nullptr;  // per-workgroup centroid bounds
    unsigned* bounds = nullptr;  // 6 order-preserving encoded floats, depth, spare, four-wide node count
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    void* scan_tmp = nullptr;  // four-wide collapse: node numbering scan
    size_t scan_tmp_bytes = 0;
    bool borrowed = false;  // buffers laid out in memory the caller owns (lbvh_bind)
};

int lbvh_reserve(LbvhWork& w, int n, std::string& err);
void lbvh_free(LbvhWork& w);
// Lays the buffers lbvh_reserve would allocate for n triangles (or cells)
// out in the caller's region at base, 256-byte aligned, and returns the
// bytes they take (base == nullptr: the size only, w untouched).  A bound
// work area is never freed or grown: builds of more than n fail.
size_t lbvh_bind(LbvhWork& w, int n, char* base, std::string& err);
// Builds the BVH of ntri triangles (vertex indices into vert, 3 floats per
// vertex).  out.nodes / out.tris must hold max(1, ntri - 1) nodes and 3 * ntri
// float4.  Enqueued on stream; depth is read back only when want_depth.
// morton_mode 0: 3D codes in a cube of the largest extent, 1: 2D x/y codes;
// width 2: binary BvhNode tree, 4: four-wide Bvh4Node tree (its node count
// is left on the device in *out.d_num_nodes and num_nodes is set to -1)
// sah_top (binary trees of 4096 or more triangles): the top of the Morton
// tree (nodes over more than max(leaf_size, 8, ntri / 8192) triangles) is
// rebuilt by a binned SAH over the subtrees below it; synchronous on stream.
// The depth of want_depth is then that of the Morton tree.
int lbvh_build(LbvhWork& w, const float* vert, const int* tri, int ntri, int leaf_size,
               int morton_mode, int width, hipStream_t stream, DeviceBvh& out, bool want_depth,
               std::string& err, bool sah_top = false);

// Order of the cells: those next to a mesh hole first (CellFrame flag bit 0),
// each group in Morton order of the ray origins (2D, x/y of the ENU frame).
int order_cells(LbvhWork& w, const CellFrame* cells, int ncell, int* order, hipStream_t stream,
                std::string& err);
int order_reserve(LbvhWork& w, int ncell, std::string& err);

// ---------------------------------------------------------------------------
// Trace kernels (trace.hip)
// ---------------------------------------------------------------------------

struct TraceCounters {
    unsigned long long rays;
    unsigned long long nodes;
    unsigned long long tris;
    unsigned long long error;
    // counters mode, persistent kernel: shader clock cycles each wavefront
    // spent refilling lanes and traversing, and traversal iterations
    unsigned long long cyc_refill;
    unsigned long long cyc_trav;
    unsigned long long trav_iters;
    // cooperative walk: queries cast ahead of the walk that it did not need
    // (rays counts the walk's own queries only)
    unsigned long long spec_rays;
};

// Device-side state of one trace launch: counters plus the work queue of
// the persistent kernel (one cell counter per XCD, 128 bytes apart).
constexpr int kQueueStride = 32;
struct TraceState {
    TraceCounters cnt;
    int pad[16];
    int queue[8 * kQueueStride];
};
static_assert(sizeof(TraceCounters) + 16 * sizeof(int) == 128, "queue offset");

// kernel: 0 = persistent lanes with per-XCD work queues (default),
//         1 = one lane per cell, azimuths in lock step
// refill: the persistent kernel leaves its traversal loop to refill lanes
//         once fewer than `refill` of the 64 lanes hold a live ray
// One domain of a batched launch (device table entry).
struct BatchDomain {
    const BvhNode* nodes;  // Bvh4Node records when the batch runs four-wide trees
    const float4* tris;
    const CellFrame* cells;
    const int* order;
    double* horizon;
    double* skyview;
    int num_tri;
    int ncell;
    int start;  // global work index of the domain's first cell
    int pad;
};

// One persistent launch over the cells of ndom domains sharing the search
// parameters p (p.ncell unused); total = sum of their cells.  mode 0 or 1.
int launch_trace_batch(const BatchDomain* doms, int ndom, int total, bool multi, int width,
                       const WalkParams& p, int mode, int block, int refill, int chunk,
                       TraceState* ts, double* spill, hipStream_t st, std::string& err);
// lanes of the batched launch (also settles its cached occupancy query, which
// must not run inside a stream capture)
int batch_grid(int mode, int block, bool multi, int width, int total);

// spill: per-lane FP64 search state of the persistent kernel, at least
// trace_spill_doubles(work items) doubles.  hitem (modes 1 and 2, persistent
// kernel; may be null): ncell * num_hori * refine_factor doubles; the
// persistent kernel then takes (cell, azimuth) work items, and the spill must
// hold trace_spill_doubles(ncell * num_hori * refine_factor) doubles.
// items_finish: with hitem, form the sector means and SVF from the item
// results (k_items_finish); false when the certificate does it instead.
// Per cell, the node its queries can start at in the binary tree bvh: from
// the root, the child whose x/y box holds the cell's ray origin while the
// other child's x/y box lies farther than reach (no ray of length tfar from
// the origin can reach a triangle there); the node's reference (index and
// wide bit) goes to start[cell].
void launch_start_nodes(const DeviceBvh& bvh, const CellFrame* cells, int ncell, float reach,
                        int* start, hipStream_t stream);
int launch_trace(const DeviceBvh& bvh, const CellFrame* cells, const int* order,
                 const WalkParams& p, int mode, int kernel, int block_size, int refill,
                 bool counters, double* horizon, double* skyview, TraceState* state,
                 double* spill, double* hitem, bool items_finish, hipStream_t stream,
                 std::string& err);
// The reference walk of the cells at the positions p.range names (at most
// max_cells), each by a team of `team` lanes (2..64, a power of two) that
// casts the walk's next steps together: the first round of an azimuth casts
// its first query with up- and down-steps beside it, later rounds the next
// `team` steps of the run, and the walk takes the first answer that ends its
// run.  Rays, elevations and results equal the one-lane walk's bit for bit;
// state->cnt.rays counts the walk's own queries, cnt.spec_rays the ones cast
// ahead and not used.  The teams take cells from one counter of the state
// (its work queues are cleared first); at most grid_blocks workgroups of 128
// lanes.  Binary trees only.
int launch_walk_team(const DeviceBvh& bvh, const CellFrame* cells, const int* order,
                     const WalkParams& p, int team, bool counters, double* horizon,
                     double* skyview, TraceState* state, int max_cells, int grid_blocks,
                     hipStream_t stream, std::string& err);
// Zero the counters and work queues of a trace state.
void launch_clear_state(TraceState* state, hipStream_t stream);
size_t trace_spill_doubles(int nwork);

// Rate (steps per second) of dependent random 64-byte reads from num_nodes
// 128-byte records at the trace kernel's occupancy (calib.hip); hot_nodes of
// them take every other step.  Allocates and frees its own device memory.
int gather_chase_rate(long long num_nodes, long long hot_nodes, int iters, double* rate,
                      std::string& err);

// Error reporting and argument checks of the C-ABI (capi.cpp): api_fail sets
// the message hz_last_error returns and passes code through.
int api_fail(int code, const std::string& msg);
void api_clear_error();
int api_check_run_args(int num_hori, double dist_search, int refine_factor, int svf_type);
// Persistent kernel: lanes are refilled once fewer than this many of 64 hold
// a live query (C2 walk: 26.8 ms at 36 against 27.3 at 32 and 27.7 at 40)
constexpr int kDefaultRefill = 36;
constexpr int S_FIELDS_HOST = 13;  // 8-byte fields of the per-lane search spill

}  // namespace hz

#define HZ_HIP(call, errstr)                                                            \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            errstr = std::string(#call) + ": " + hipGetErrorString(e_);                 \
            return -2;                                                                  \
        }                                                                               \
    } while (0)
