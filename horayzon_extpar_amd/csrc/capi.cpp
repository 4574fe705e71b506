// C-ABI of libhorayzon.so (declared in include/horayzon.h).
//
// horizon_svf_comp keeps the reference signature and output layout
// (mo_lradtopo_horayzon.h:4-11, mo_lradtopo_horayzon.cpp:562-785).  The
// reference has no C-side validation and no status; this layer validates the
// same ranges the reference's Python binding checks (horizon_svf.pyx:79-103),
// returns a status from the _ex / hz_domain_* entry points and fills the
// outputs with NaN on failure.  Global mutable state: the per-thread error
// message, and the one-shot calls' buffer cache, whose mutex hands a kept
// domain to one call at a time; calls are reentrant (each call owns its
// stream and buffers while it runs).

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

#include "../../include/horayzon.h"
#include "hz_internal.h"
#include "learn_internal.h"
#include "mlp.h"

using namespace hz;

namespace {

thread_local std::string g_err;

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

bool verbose_env() {
    const char* v = std::getenv("HZ_VERBOSE");
    return v && v[0] && v[0] != '0';
}

void fill_nan(double* horizon, double* skyview, int num_hori, int num_cell) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (horizon && num_hori > 0 && num_cell > 0)
        for (size_t i = 0; i < (size_t)num_hori * (size_t)num_cell; i++) horizon[i] = nan;
    if (skyview && num_cell > 0)
        for (int i = 0; i < num_cell; i++) skyview[i] = nan;
}

int check_grid_args(const int* cov, int num_cell, int num_vertex, int grid_type,
                    double ray_org_elev) {
    if (num_cell < 1) return fail(HZ_EINVAL, "num_cell must be >= 1");
    if (num_vertex < 0) return fail(HZ_EINVAL, "num_vertex must be >= 0");
    if (grid_type < 0 || grid_type > 1)
        return fail(HZ_EINVAL, "'grid_type' must be in the range [0, 1]");
    if (!(ray_org_elev >= 0.1)) return fail(HZ_EINVAL, "Minimal allowed value for 'ray_org_elev' is 0.1 m");
    for (size_t k = 0; k < (size_t)6 * (size_t)num_vertex; k++) {
        int c = cov[k];
        if (!((c >= 0 && c <= num_cell - 1) || c == -2))
            return fail(HZ_EINVAL, "Indices of 'cells_of_vertex' out of range");
    }
    return 0;
}

int check_run_args(int num_hori, double dist_search, int refine_factor, int svf_type) {
    if (num_hori < 4 || num_hori > 1440)
        return fail(HZ_EINVAL, "'num_hori' must be in the range [4, 1440]");
    if (!(dist_search >= 1000.0 && dist_search <= 500000.0))
        return fail(HZ_EINVAL, "'dist_search' must be in the range [1'000, 500'000] m");
    if (refine_factor < 1 || refine_factor > 50)
        return fail(HZ_EINVAL, "'refine_factor' must be in the range [1, 50]");
    if (svf_type < 0 || svf_type > 2)
        return fail(HZ_EINVAL, "'svf_type' must be in the range [0, 2]");
    return 0;
}

}  // namespace

namespace hz {
int learn_fail(int code, const std::string& msg) { return fail(code, msg); }
int api_fail(int code, const std::string& msg) { return fail(code, msg); }
void api_clear_error() { g_err.clear(); }
int api_check_run_args(int num_hori, double dist_search, int refine_factor, int svf_type) {
    return check_run_args(num_hori, dist_search, refine_factor, svf_type);
}
}  // namespace hz

struct hz_domain {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncell = 0;
    int grid_type = 1;
    long long ntri = 0;
    long long nmeshvert = 0;
    double lon0 = 0, lat0 = 0;
    float* d_vert = nullptr;
    int* d_tri = nullptr;
    CellFrame* d_cells = nullptr;
    int* d_order = nullptr;
    int* d_start = nullptr;  // per cell: the node its queries start at (binary trees)
    BvhNode* d_nodes = nullptr;
    float4* d_tris = nullptr;
    double* d_horizon = nullptr;
    double* d_skyview = nullptr;
    int out_nhori = 0;  // horizon rows held by d_horizon
    int last_nhori = 0; // rows written by the last run
    TraceState* d_state = nullptr;
    double* d_spill = nullptr;
    size_t spill_doubles = 0;  // capacity of d_spill
    double* d_hitem = nullptr;  // per-(azimuth, cell) results of the lattice searches
    size_t hitem_doubles = 0;
    double2* d_elev_cs = nullptr;
    int* d_adj_off = nullptr;  // cell -> adjacent triangles (slope / aspect)
    int* d_adj = nullptr;
    double* d_slope = nullptr;
    Stager stager;  // pinned slots of the large uploads and downloads
    float bounds_xy[4] = {0, 0, 0, 0};  // scene x/y extent (raster of the descriptor)
    // learned mode
    const hz_predictor* predictor = nullptr;
    Raster raster;
    uint16_t* d_desc = nullptr;
    float* d_pred = nullptr;
    bool have_pred = false;
    LbvhWork work;
    DeviceBvh bvh;
    // certified lattice searches: mesh-boundary index, per-item bounds, the
    // cells handed to the walk, and the cell elevation range
    CertIndex cert;
    double hmin = 0.0, hmax = 0.0;
    // certified lattice searches: cells next to mesh holes are walked first
    hipEvent_t bev[2] = {nullptr, nullptr};  // start of the three launches, border walk done
    TraceState* d_state2 = nullptr;  // the border walk
    TraceState* d_state3 = nullptr;  // the far cells' items
    hipEvent_t cev[4] = {nullptr, nullptr, nullptr, nullptr};  // near walk / far items done
    int* d_ranges = nullptr;  // class counts and {first, count} position ranges
    bool near_flags = false;  // flag bit 1 set on the cells (certified runs)
    hipStream_t xstream = nullptr;  // far cells' lattice items
    hipStream_t ystream = nullptr;  // border cells' team walk
    double* d_spill_x = nullptr;    // spill area of the far cells' items
    size_t spill_x_doubles = 0;
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // learned search: the predictor stage runs on its own stream beside the
    // LBVH build and the cell ordering (created on first use)
    hipStream_t pstream = nullptr;
    hipEvent_t pev[3] = {nullptr, nullptr, nullptr};  // fork, predictor start, join
    bool ran = false;
    bool host_scene = false;  // built from a host prologue (the one-shot calls' cache holds these)
    // a kept domain reloaded by a one-shot call builds its default tree while
    // the host computes the cell frames; the run then skips the build
    bool bvh_pre = false;
    float bvh_pre_ms = 0.0f;
    hipEvent_t pre_ev[2] = {nullptr, nullptr};
    double* d_pcell = nullptr;  // FP64 ENU cell points of a reload (frames built on the device)
    int* d_holes = nullptr;     // cells next to mesh holes of a reload
    size_t holes_cap = 0;
};

// The progress report the reference prints to stdout (mo_lradtopo_horayzon.cpp
// :240-315, :427, :589-680, :765-776), as text: mesh, ENU origin, SVF form,
// BVH build, tracing and the ray statistics (rays per cell and output sector,
// :770), or the failure.
static std::string run_report(const hz_stats& s, int num_cell, int num_hori, int grid_type,
                              int svf_type, int rc) {
    const double r2d = 180.0 / 3.14159265358979323846;
    char line[512];
    std::string out;
    const char* rule =
        "-------------------------------------------------------------------------------\n";
    out += rule;
    out += "Horizon and SVF computation with HIP on gfx950\n";
    out += rule;
    if (rc != 0 && s.num_triangles == 0) {
        std::snprintf(line, sizeof(line), "horizon_svf_comp failed (%d): %s\n", rc, g_err.c_str());
        return out + line;
    }
    out += grid_type == 0 ? "Triangle mesh from ICON grid cell circumcenters only\n"
                          : "Triangle mesh from ICON grid cell circumcenters and vertices\n";
    std::snprintf(line, sizeof(line),
                  "Number of triangles in mesh: %lld\n"
                  "Triangle mesh building and coordinate transforms: %.2f s\n"
                  "Origin of ENU coordinate system: %.4f deg lat, %.4f deg lon\n",
                  s.num_triangles, s.t_prologue_ms * 1e-3, s.lat0 * r2d, s.lon0 * r2d);
    out += line;
    std::snprintf(line, sizeof(line), "Sky view factor: %s\n",
                  svf_type == 0   ? "pure geometric SVF"
                  : svf_type == 1 ? "geometric scaled with sin(horizon)"
                                  : "geometric scaled with sin(horizon)**2");
    out += line;
    std::snprintf(line, sizeof(line),
                  "Building bounding volume hierarchy (BVH): %.2f s\n"
                  "Ray tracing: %.2f s\n"
                  "Number of rays shot: %lld\n"
                  "Average number of rays per cell and azimuth sector: %.2f\n",
                  s.t_bvh_ms * 1e-3, s.t_trace_ms * 1e-3, s.num_rays,
                  num_cell > 0 && num_hori > 0 ? (double)s.num_rays / ((double)num_cell * num_hori)
                                               : 0.0);
    out += line;
    if (rc != 0) {
        std::snprintf(line, sizeof(line), "horizon_svf_comp failed (%d): %s\n", rc,
                      g_err.c_str());
        out += line;
    }
    out += rule;
    return out;
}

extern "C" {

const char* hz_last_error(void) { return g_err.c_str(); }

void hz_default_options(hz_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->search_mode = HZ_MODE_WALK;
    o->leaf_size = 1;
    o->block_size = 128;
    o->collect_counters = 0;
    o->cell_order = 1;
    o->verbose = 0;
    o->trace_kernel = 0;
    o->refill = kDefaultRefill;
    o->queue_chunk = 0;
    o->bvh_width = 2;
    o->prologue = HZ_PROLOGUE_HOST;
    o->whole_cells = 0;
    o->certify = 1;
    o->cache = 1;
}

void hz_domain_destroy(hz_domain* d) {
    if (!d) return;
    (void)hipSetDevice(d->device);
    if (d->pstream) (void)hipStreamSynchronize(d->pstream);
    if (d->xstream) (void)hipStreamSynchronize(d->xstream);
    if (d->ystream) (void)hipStreamSynchronize(d->ystream);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    void* bufs[] = {d->d_vert,   d->d_tri,     d->d_cells,   d->d_order, d->d_nodes,
                    d->d_tris,   d->d_horizon, d->d_skyview, d->d_state,
                    d->d_spill,  d->d_elev_cs, d->d_desc,    d->d_pred,
                    d->d_adj_off, d->d_adj,    d->d_slope,   d->d_hitem,
                    d->d_state2, d->d_state3, d->d_ranges, d->d_spill_x,
                    d->d_pcell,  d->d_holes,  d->d_start};
    d->stager.release();
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    lbvh_free(d->work);
    cert_index_free(d->cert);
    d->raster.free_all();
    for (auto& e : d->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : d->pev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : d->bev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : d->cev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : d->pre_ev)
        if (e) (void)hipEventDestroy(e);
    if (d->xstream) (void)hipStreamDestroy(d->xstream);
    if (d->ystream) (void)hipStreamDestroy(d->ystream);

    if (d->pstream) (void)hipStreamDestroy(d->pstream);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

int hz_domain_num_cell(const hz_domain* d) { return d ? d->ncell : 0; }

int hz_scene_export(const double* clon, const double* clat, const double* hsurf,
                    const double* vlon, const double* vlat, const int* cov, int num_cell,
                    int num_vertex, int grid_type, double ray_org_elev, long long* num_triangles,
                    long long* num_mesh_vertices, float* vert32, int* tri, double* frame_n,
                    double* frame_north, float* org, double* lon0_lat0) {
    g_err.clear();
    if (!clon || !clat || !hsurf || (num_vertex > 0 && (!vlon || !vlat || !cov)))
        return fail(HZ_EINVAL, "NULL input array");
    int rc = check_grid_args(cov, num_cell, num_vertex, grid_type, ray_org_elev);
    if (rc) return rc;
    HostScene hs;
    std::string err;
    if (prologue(clon, clat, hsurf, vlon, vlat, cov, num_cell, num_vertex, grid_type,
                 ray_org_elev, hs, err))
        return fail(HZ_EINVAL, err);
    if (num_triangles) *num_triangles = (long long)hs.tri.size() / 3;
    if (num_mesh_vertices) *num_mesh_vertices = hs.num_mesh_vertices;
    if (vert32) std::memcpy(vert32, hs.vert32.data(), hs.vert32.size() * sizeof(float));
    if (tri) std::memcpy(tri, hs.tri.data(), hs.tri.size() * sizeof(int));
    for (int i = 0; i < num_cell; i++) {
        const CellFrame& f = hs.frame[i];
        for (int a = 0; a < 3; a++) {
            if (frame_n) frame_n[3 * i + a] = f.n[a];
            if (frame_north) frame_north[3 * i + a] = f.north[a];
            if (org) org[3 * i + a] = f.org[a];
        }
    }
    if (lon0_lat0) {
        lon0_lat0[0] = hs.lon0;
        lon0_lat0[1] = hs.lat0;
    }
    return HZ_OK;
}
long long hz_domain_num_triangles(const hz_domain* d) { return d ? d->ntri : 0; }

#define DCHK(call)                                                                     \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            hz_domain_destroy(d);                                                      \
            return fail(e_ == hipErrorOutOfMemory ? HZ_ENOMEM : HZ_EDEVICE,           \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
        }                                                                              \
    } while (0)

static int make_domain(const double* clon, const double* clat, const double* hsurf,
                       const double* vlon, const double* vlat, const int* cov, int num_cell,
                       int num_vertex, int grid_type, double ray_org_elev, int where,
                       const HostScene& hs, double t0, hz_domain** out, hz_stats* stats);

int hz_domain_create(const double* clon, const double* clat, const double* hsurf,
                     const double* vlon, const double* vlat, const int* cov, int num_cell,
                     int num_vertex, int grid_type, double ray_org_elev, const hz_options* opt,
                     hz_domain** out, hz_stats* stats) {
    g_err.clear();
    if (!out) return fail(HZ_EINVAL, "out is NULL");
    *out = nullptr;
    if (!clon || !clat || !hsurf || (num_vertex > 0 && (!vlon || !vlat || !cov)))
        return fail(HZ_EINVAL, "NULL input array");
    int rc = check_grid_args(cov, num_cell, num_vertex, grid_type, ray_org_elev);
    if (rc) return rc;
    const int where = opt ? opt->prologue : HZ_PROLOGUE_HOST;
    if (where != HZ_PROLOGUE_HOST && where != HZ_PROLOGUE_DEVICE)
        return fail(HZ_EINVAL, "prologue must be HZ_PROLOGUE_HOST or HZ_PROLOGUE_DEVICE");
    double t0 = now_ms();
    std::string err;
    HostScene hs;
    if (where == HZ_PROLOGUE_HOST) {
        rc = prologue(clon, clat, hsurf, vlon, vlat, cov, num_cell, num_vertex, grid_type,
                      ray_org_elev, hs, err);
        if (rc) return fail(HZ_EINVAL, err);
    }
    return make_domain(clon, clat, hsurf, vlon, vlat, cov, num_cell, num_vertex, grid_type,
                       ray_org_elev, where, hs, t0, out, stats);
}

// The device side of hz_domain_create: buffers, streams and events, then the
// host scene's upload (or the device prologue).
static int make_domain(const double* clon, const double* clat, const double* hsurf,
                       const double* vlon, const double* vlat, const int* cov, int num_cell,
                       int num_vertex, int grid_type, double ray_org_elev, int where,
                       const HostScene& hs, double t0, hz_domain** out, hz_stats* stats) {
    int rc = 0;
    std::string err;
    double t1 = now_ms();
    hz_domain* d = new hz_domain();
    {
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || ndev < 1) {
            delete d;
            return fail(HZ_EDEVICE, "no HIP device available");
        }
    }
    DCHK(hipGetDevice(&d->device));
    DCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    for (auto& e : d->ev) DCHK(hipEventCreate(&e));
    d->ncell = num_cell;
    d->grid_type = grid_type;
    d->hmin = d->hmax = hsurf[0];
    for (int i = 1; i < num_cell; i++) {
        d->hmin = std::min(d->hmin, hsurf[i]);
        d->hmax = std::max(d->hmax, hsurf[i]);
    }
    if (where == HZ_PROLOGUE_DEVICE) {
        DeviceScene ds;
        rc = prologue_device(clon, clat, hsurf, vlon, vlat, cov, num_cell, num_vertex, grid_type,
                             ray_org_elev, d->stream, d->stager, ds, err);
        if (rc) {
            hz_domain_destroy(d);
            return fail(rc == -1 ? HZ_EINVAL : (rc == -3 ? HZ_ENOMEM : HZ_EDEVICE), err);
        }
        d->d_vert = ds.vert;
        d->d_tri = ds.tri;
        d->d_cells = ds.cells;
        d->ntri = ds.ntri;
        d->nmeshvert = ds.nmv;
        d->lon0 = ds.lon0;
        d->lat0 = ds.lat0;
        std::memcpy(d->bounds_xy, ds.bounds_xy, sizeof(d->bounds_xy));
        t1 = now_ms();
    } else {
        d->host_scene = true;
        d->ntri = (long long)hs.tri.size() / 3;
        d->nmeshvert = hs.num_mesh_vertices;
        d->lon0 = hs.lon0;
        d->lat0 = hs.lat0;
        size_t nv = hs.vert32.size(), nt = hs.tri.size();
        DCHK(hipMalloc(&d->d_vert, std::max<size_t>(nv, 3) * sizeof(float)));
        DCHK(hipMalloc(&d->d_tri, std::max<size_t>(nt, 3) * sizeof(int)));
        DCHK(hipMalloc(&d->d_cells, (size_t)num_cell * sizeof(CellFrame)));
        std::memcpy(d->bounds_xy, hs.bounds_xy, sizeof(d->bounds_xy));
        if (nv) DCHK(d->stager.copy(d->d_vert, hs.vert32.data(), nv * sizeof(float), true, d->stream));
        if (nt) DCHK(d->stager.copy(d->d_tri, hs.tri.data(), nt * sizeof(int), true, d->stream));
        DCHK(d->stager.copy(d->d_cells, hs.frame.data(), (size_t)num_cell * sizeof(CellFrame), true,
                            d->stream));
    }
    DCHK(hipMalloc(&d->d_order, (size_t)num_cell * sizeof(int)));
    DCHK(hipMalloc(&d->d_nodes, std::max<long long>(1, d->ntri - 1) * sizeof(BvhNode)));
    DCHK(hipMalloc(&d->d_tris, std::max<long long>(1, d->ntri) * 3 * sizeof(float4)));
    DCHK(hipMalloc(&d->d_skyview, (size_t)num_cell * sizeof(double)));
    DCHK(hipMalloc(&d->d_state, sizeof(TraceState)));
    d->spill_doubles = trace_spill_doubles(num_cell);
    DCHK(hipMalloc(&d->d_spill, d->spill_doubles * sizeof(double)));
    DCHK(hipMalloc(&d->d_elev_cs, (kMaxLattice + 1) * (sizeof(double2) + sizeof(double))));
    {
        std::vector<double2> cs(kMaxLattice + 1);
        std::vector<double> sh(kMaxLattice + 1);
        lattice_table(make_walk_params(24, 1, 2, 4e4, num_cell).hori_acc, cs.data(), sh.data());
        DCHK(hipMemcpy(d->d_elev_cs, cs.data(), cs.size() * sizeof(double2), hipMemcpyHostToDevice));
        DCHK(hipMemcpy(d->d_elev_cs + cs.size(), sh.data(), sh.size() * sizeof(double),
                       hipMemcpyHostToDevice));
    }
    if (lbvh_reserve(d->work, (int)std::max<long long>(std::max<long long>(d->ntri, num_cell), 1), err)) {
        hz_domain_destroy(d);
        return fail(HZ_ENOMEM, err);
    }
    DCHK(hipStreamSynchronize(d->stream));
    double t2 = now_ms();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->num_triangles = d->ntri;
        stats->num_mesh_vertices = d->nmeshvert;
        stats->t_prologue_ms = t1 - t0;
        stats->t_upload_ms = t2 - t1;
        stats->lon0 = d->lon0;
        stats->lat0 = d->lat0;
    }
    *out = d;
    return HZ_OK;
}

int hz_domain_scene(const hz_domain* d, float* vert32, int* tri, double* frame_n,
                    double* frame_north, float* org) {
    g_err.clear();
    if (!d) return fail(HZ_EINVAL, "domain is NULL");
    hipError_t e = hipSuccess;
    if (vert32 && d->nmeshvert)
        e = hipMemcpy(vert32, d->d_vert, (size_t)d->nmeshvert * 3 * sizeof(float),
                      hipMemcpyDeviceToHost);
    if (e == hipSuccess && tri && d->ntri)
        e = hipMemcpy(tri, d->d_tri, (size_t)d->ntri * 3 * sizeof(int), hipMemcpyDeviceToHost);
    std::vector<CellFrame> f;
    if (e == hipSuccess && (frame_n || frame_north || org)) {
        f.resize(d->ncell);
        e = hipMemcpy(f.data(), d->d_cells, (size_t)d->ncell * sizeof(CellFrame),
                      hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return fail(HZ_EDEVICE, std::string("hz_domain_scene: ") + hipGetErrorString(e));
    for (size_t i = 0; i < f.size(); i++)
        for (int a = 0; a < 3; a++) {
            if (frame_n) frame_n[3 * i + a] = f[i].n[a];
            if (frame_north) frame_north[3 * i + a] = f[i].north[a];
            if (org) org[3 * i + a] = f[i].org[a];
        }
    return HZ_OK;
}

int hz_calibrate_gather(long long num_nodes, long long hot_nodes, int iters, double* rate) {
    g_err.clear();
    std::string err;
    const int rc = gather_chase_rate(num_nodes, hot_nodes, iters, rate, err);
    if (rc == -1) return fail(HZ_EINVAL, err);
    if (rc) return fail(HZ_EDEVICE, err);
    return HZ_OK;
}

int hz_sincos(const double* x, long long n, double* s, double* c, int on_device) {
    g_err.clear();
    if (n < 0 || (n > 0 && (!x || !s || !c))) return fail(HZ_EINVAL, "hz_sincos: bad arguments");
    std::string err;
    if (sincos_cr_eval(x, n, s, c, on_device, err)) return fail(HZ_EDEVICE, err);
    return HZ_OK;
}

#undef DCHK
#define RCHK(call)                                                                     \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? HZ_ENOMEM : HZ_EDEVICE,           \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

// Raster, descriptors and predictor on all cells of d (enqueued on st).
static int predict_cells(hz_domain* d, hipStream_t st, std::string& err) {
    const int n = d->ncell;
    if (!d->d_desc) {
        if (hipMalloc(&d->d_desc, (size_t)n * kIn * sizeof(uint16_t)) != hipSuccess ||
            hipMalloc(&d->d_pred, (size_t)n * kOutPad * sizeof(float)) != hipSuccess) {
            err = "hipMalloc failed (descriptors)";
            return HZ_ENOMEM;
        }
    }
    const float* b = d->bounds_xy;
    double area = std::max(1.0, (double)(b[2] - b[0]) * (double)(b[3] - b[1]));
    float pix = (float)std::sqrt(area / std::max<long long>(1, d->ntri));
    if (raster_build(d->raster, d->d_vert, d->d_tri, (int)d->ntri, b, pix, st, err))
        return HZ_EDEVICE;
    descriptors(d->raster, d->d_cells, n, d->d_desc, st);
    if (d->predictor) {
        mlp_infer(d->predictor->net, d->d_desc, n, d->d_pred, st);
        d->have_pred = true;
    }
    if (hipGetLastError() != hipSuccess) {
        err = "descriptor / predictor launch failed";
        return HZ_EDEVICE;
    }
    return 0;
}

int hz_domain_set_predictor(hz_domain* d, const hz_predictor* p) {
    g_err.clear();
    if (!d) return fail(HZ_EINVAL, "domain is NULL");
    d->predictor = p;
    d->have_pred = false;
    return HZ_OK;
}

int hz_domain_descriptors(hz_domain* d, unsigned short* desc) {
    g_err.clear();
    if (!d || !desc) return fail(HZ_EINVAL, "NULL argument");
    RCHK(hipSetDevice(d->device));
    std::string err;
    const hz_predictor* keep = d->predictor;
    d->predictor = nullptr;  // descriptors only
    int rc = predict_cells(d, d->stream, err);
    d->predictor = keep;
    if (rc) return fail(rc, err);
    RCHK(hipMemcpyAsync(desc, d->d_desc, (size_t)d->ncell * kIn * sizeof(uint16_t),
                        hipMemcpyDeviceToHost, d->stream));
    RCHK(hipStreamSynchronize(d->stream));
    return HZ_OK;
}

int hz_domain_slope_aspect(hz_domain* d, double* slope, double* aspect) {
    g_err.clear();
    if (!d || !slope || !aspect) return fail(HZ_EINVAL, "NULL argument");
    RCHK(hipSetDevice(d->device));
    if (!d->d_adj_off) {  // cell -> triangle adjacency, built on first use
        std::vector<int> tri((size_t)d->ntri * 3), off, adj;
        if (d->ntri)
            RCHK(hipMemcpyAsync(tri.data(), d->d_tri, tri.size() * sizeof(int),
                                hipMemcpyDeviceToHost, d->stream));
        RCHK(hipStreamSynchronize(d->stream));
        cell_adjacency(tri.data(), d->ntri, d->ncell, off, adj);
        if (d->d_adj) RCHK(hipFree(d->d_adj));  // left by an earlier failed attempt
        d->d_adj = nullptr;
        RCHK(hipMalloc(&d->d_adj, std::max<size_t>(adj.size(), 1) * sizeof(int)));
        if (!adj.empty())
            RCHK(hipMemcpy(d->d_adj, adj.data(), adj.size() * sizeof(int), hipMemcpyHostToDevice));
        RCHK(hipMalloc(&d->d_adj_off, off.size() * sizeof(int)));
        RCHK(hipMemcpy(d->d_adj_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    if (!d->d_slope) RCHK(hipMalloc(&d->d_slope, (size_t)d->ncell * 2 * sizeof(double)));
    slope_aspect(d->d_vert, d->d_tri, d->d_adj_off, d->d_adj, d->d_cells, d->ncell, d->d_slope,
                 d->d_slope + d->ncell, d->stream);
    RCHK(hipGetLastError());
    RCHK(hipMemcpyAsync(slope, d->d_slope, (size_t)d->ncell * sizeof(double),
                        hipMemcpyDeviceToHost, d->stream));
    RCHK(hipMemcpyAsync(aspect, d->d_slope + d->ncell, (size_t)d->ncell * sizeof(double),
                        hipMemcpyDeviceToHost, d->stream));
    RCHK(hipStreamSynchronize(d->stream));
    return HZ_OK;
}

int hz_domain_predictions(hz_domain* d, float* pred) {
    g_err.clear();
    if (!d || !pred) return fail(HZ_EINVAL, "NULL argument");
    if (!d->have_pred) return fail(HZ_EINVAL, "no learned run to read predictions from");
    RCHK(hipSetDevice(d->device));
    std::vector<float> tmp((size_t)d->ncell * kOutPad);
    RCHK(hipMemcpyAsync(tmp.data(), d->d_pred, tmp.size() * sizeof(float), hipMemcpyDeviceToHost,
                        d->stream));
    RCHK(hipStreamSynchronize(d->stream));
    for (int i = 0; i < d->ncell; i++)
        std::memcpy(pred + (size_t)i * kOut, tmp.data() + (size_t)i * kOutPad, kOut * sizeof(float));
    return HZ_OK;
}

int hz_domain_run(hz_domain* d, int num_hori, double dist_search, int refine_factor,
                  int svf_type, const hz_options* opt, hz_stats* stats) {
    g_err.clear();
    if (!d) return fail(HZ_EINVAL, "domain is NULL");
    int rc = check_run_args(num_hori, dist_search, refine_factor, svf_type);
    if (rc) return rc;
    hz_options o;
    if (opt)
        o = *opt;
    else
        hz_default_options(&o);
    if (o.search_mode != HZ_MODE_WALK && o.search_mode != HZ_MODE_CLIMB &&
        o.search_mode != HZ_MODE_LEARNED)
        return fail(HZ_EINVAL, "unknown search_mode");
    if (o.search_mode == HZ_MODE_LEARNED) {
        if (!d->predictor) return fail(HZ_EINVAL, "learned search: no predictor set");
        if (num_hori * refine_factor != kDescAz)
            return fail(HZ_EINVAL, "learned search needs num_hori * refine_factor == 24");
        if (o.trace_kernel == 1)
            return fail(HZ_EINVAL, "learned search needs a persistent trace kernel");
    }
    int block = o.block_size ? o.block_size : 128;
    if (block != 64 && block != 128 && block != 256)
        return fail(HZ_EINVAL, "block_size must be 64, 128 or 256");
    int leaf = o.leaf_size ? o.leaf_size : 1;
    if (leaf < 1 || leaf > kMaxLeafSize) return fail(HZ_EINVAL, "leaf_size must be in [1, 8]");
    int width = o.bvh_width ? o.bvh_width : 2;
    if (width != 2 && width != 4) return fail(HZ_EINVAL, "bvh_width must be 2 or 4");
    if (o.bvh_morton < 0 || o.bvh_morton > 1) return fail(HZ_EINVAL, "bvh_morton must be 0 or 1");
    if (o.trace_kernel < 0 || o.trace_kernel > 1)
        return fail(HZ_EINVAL, "trace_kernel must be 0 or 1");
    int refill = o.refill ? o.refill : kDefaultRefill;
    if (refill < 1 || refill > 64) return fail(HZ_EINVAL, "refill must be in [1, 64]");
    if (o.trace_kernel == 1) width = 2;  // the lane-per-cell kernel walks binary trees
    if (o.certify < 0 || o.certify > 1) return fail(HZ_EINVAL, "certify must be 0 or 1");
    if (o.bvh_sah < 0 || o.bvh_sah > 1) return fail(HZ_EINVAL, "bvh_sah must be 0 or 1");
    if (o.bvh_sah && width != 2) return fail(HZ_EINVAL, "bvh_sah needs the binary tree (bvh_width 2)");
    const bool certify = o.certify == 1 && o.search_mode != HZ_MODE_WALK;
    if (certify && (o.trace_kernel == 1 || o.whole_cells))
        return fail(HZ_EINVAL,
                    "certify needs (cell, azimuth) work items: persistent kernel, whole_cells 0 "
                    "(or certify 0 for the plain lattice search)");
    bool verbose = o.verbose || verbose_env();
    RCHK(hipSetDevice(d->device));
    double t0 = now_ms();
    if (num_hori > d->out_nhori) {
        if (d->d_horizon) RCHK(hipFree(d->d_horizon));
        d->d_horizon = nullptr;
        RCHK(hipMalloc(&d->d_horizon, (size_t)num_hori * d->ncell * sizeof(double)));
        d->out_nhori = num_hori;
    }
    std::string err;
    hipStream_t st = d->stream;
    RCHK(hipEventRecord(d->ev[0], st));
    const bool learned = o.search_mode == HZ_MODE_LEARNED;
    if (learned) {
        // the predictor stage (raster, descriptors, network) reads only the
        // scene and the frames: it runs on a stream of its own beside the
        // LBVH build and the cell ordering, and the trace waits for both
        if (!d->pstream) {
            hipStream_t s2 = nullptr;
            hipEvent_t e[3] = {nullptr, nullptr, nullptr};
            hipError_t e_ = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking);
            for (int i = 0; i < 3 && e_ == hipSuccess; i++) e_ = hipEventCreate(&e[i]);
            if (e_ != hipSuccess) {
                for (auto& x : e)
                    if (x) (void)hipEventDestroy(x);
                if (s2) (void)hipStreamDestroy(s2);
                return fail(HZ_EDEVICE, std::string("predictor stream: ") + hipGetErrorString(e_));
            }
            d->pstream = s2;
            for (int i = 0; i < 3; i++) d->pev[i] = e[i];
        }
        RCHK(hipEventRecord(d->pev[0], st));
        RCHK(hipStreamWaitEvent(d->pstream, d->pev[0], 0));
        RCHK(hipEventRecord(d->pev[1], d->pstream));
        rc = predict_cells(d, d->pstream, err);
        if (rc) return fail(rc, err);
        RCHK(hipEventRecord(d->pev[2], d->pstream));
    }
    // the lattice searches (climb, learned start) run one work item per cell
    // and azimuth in the persistent kernel: a per-item result buffer, and the
    // search spill sized for the items' lanes
    double* hitem = nullptr;
    const long long items = (long long)d->ncell * num_hori * refine_factor;
    if (o.search_mode != HZ_MODE_WALK && o.trace_kernel == 0 && !o.whole_cells &&
        items < (1LL << 30)) {
        // without the memory for them the search runs a cell per work item
        // (same results) in the buffers the domain already holds
        bool room = true;
        if ((size_t)items > d->hitem_doubles) {
            if (d->d_hitem) RCHK(hipFree(d->d_hitem));
            d->d_hitem = nullptr;
            d->hitem_doubles = 0;
            if (hipMalloc(&d->d_hitem, (size_t)items * sizeof(double)) == hipSuccess) {
                d->hitem_doubles = (size_t)items;
            } else {
                (void)hipGetLastError();
                d->d_hitem = nullptr;
                room = false;
            }
        }
        const size_t need = trace_spill_doubles((int)items);
        if (room && need > d->spill_doubles) {
            double* sp = nullptr;
            if (hipMalloc(&sp, need * sizeof(double)) == hipSuccess) {
                RCHK(hipFree(d->d_spill));
                d->d_spill = sp;
                d->spill_doubles = need;
            } else {
                (void)hipGetLastError();
                room = false;
            }
        }
        if (room) hitem = d->d_hitem;
    }
    int mode = o.search_mode;
    if (certify && !hitem) mode = HZ_MODE_WALK;  // the walk itself: the same results
    const bool cert_run = certify && hitem;
    const float tfar = (float)dist_search;
    if (cert_run) {
        // the side streams and events of the certified searches
        if (!d->bev[0]) {
            hipError_t e_ = hipSuccess;
            for (int i = 0; i < 2 && e_ == hipSuccess; i++) e_ = hipEventCreate(&d->bev[i]);
            for (int i = 0; i < 4 && e_ == hipSuccess; i++) e_ = hipEventCreate(&d->cev[i]);
            if (e_ == hipSuccess) e_ = hipStreamCreateWithFlags(&d->xstream, hipStreamNonBlocking);
            if (e_ == hipSuccess) e_ = hipStreamCreateWithFlags(&d->ystream, hipStreamNonBlocking);
            if (e_ != hipSuccess)
                return fail(HZ_EDEVICE, std::string("certified search streams: ") +
                                            hipGetErrorString(e_));
        }
        // mesh-boundary index, then the cell classes (flag bit 1: a boundary
        // point in reach) that the trace order sorts by; the classes run on a
        // side stream beside the tree build, and the ordering waits for them
        rc = cert_index_build(d->cert, d->d_cells, d->d_vert, d->ncell, d->d_tri, d->ntri,
                              d->bounds_xy, d->hmin, d->hmax, tfar, st, err);
        if (rc) return fail(rc == -3 ? HZ_ENOMEM : HZ_EDEVICE, err);
        if (!d->d_ranges) RCHK(hipMalloc(&d->d_ranges, 10 * sizeof(int)));
        RCHK(hipEventRecord(d->cev[2], st));
        RCHK(hipStreamWaitEvent(d->xstream, d->cev[2], 0));
        RCHK(hipEventRecord(d->ev[5], d->xstream));
        launch_cert_classes(d->cert, d->d_cells, d->ncell, false, d->d_ranges, d->xstream);
        RCHK(hipGetLastError());
        RCHK(hipEventRecord(d->ev[6], d->xstream));
        d->near_flags = true;
    } else if (d->near_flags) {  // the plain order again (bit 0 first, then Morton)
        if (!d->d_ranges) RCHK(hipMalloc(&d->d_ranges, 10 * sizeof(int)));
        launch_cert_classes(d->cert, d->d_cells, d->ncell, true, d->d_ranges, st);
        RCHK(hipGetLastError());
        d->near_flags = false;
    }
    d->bvh.nodes = d->d_nodes;
    d->bvh.tris = d->d_tris;
    // the tree of the default options may already be built (one-shot reload)
    const bool use_pre = d->bvh_pre && leaf == 1 && o.bvh_morton == 0 && width == 2 &&
                         o.collect_counters == 0 && o.bvh_sah == 0;
    d->bvh_pre = false;
    if (!use_pre) {
        rc = lbvh_build(d->work, d->d_vert, d->d_tri, (int)d->ntri, leaf, o.bvh_morton, width,
                        st, d->bvh, o.collect_counters != 0, err, o.bvh_sah == 1);
        if (rc) return fail(rc == -3 ? HZ_ENOMEM : (rc == -1 ? HZ_EINVAL : HZ_EDEVICE), err);
    }
    RCHK(hipEventRecord(d->ev[1], st));
    if (cert_run) RCHK(hipStreamWaitEvent(st, d->ev[6], 0));  // the cell classes
    const int* order = nullptr;
    if (o.cell_order) {
        rc = order_cells(d->work, d->d_cells, d->ncell, d->d_order, st, err);
        if (rc) return fail(HZ_EDEVICE, err);
        order = d->d_order;
    }
    RCHK(hipEventRecord(d->ev[2], st));
    WalkParams p = make_walk_params(num_hori, refine_factor, svf_type, dist_search, d->ncell);
    p.elev_cs = d->d_elev_cs;
    p.sin_h = reinterpret_cast<const double*>(d->d_elev_cs + (kMaxLattice + 1));
    if (learned) {
        RCHK(hipStreamWaitEvent(st, d->pev[2], 0));
        p.pred = d->d_pred;
        p.pred_stride = kOutPad;
    }
    RCHK(hipEventRecord(d->ev[4], st));
    if (width == 2) {  // queries start below the levels whose other child is out of reach
        if (!d->d_start) RCHK(hipMalloc(&d->d_start, (size_t)d->ncell * sizeof(int)));
        launch_start_nodes(d->bvh, d->d_cells, d->ncell, p.tfar_box + 1.0f, d->d_start, st);
        RCHK(hipGetLastError());
        p.start = d->d_start;
    }
    bool multi_state = false;  // the border walk and far items keep states of their own
    float ms_cert = 0, ms_cwalk = 0, ms_bwalk = 0, ms_lattice = 0;
    if (!cert_run) {
        rc = launch_trace(d->bvh, d->d_cells, order, p, mode, o.trace_kernel, block, refill,
                          o.collect_counters != 0, d->d_horizon, d->d_skyview, d->d_state,
                          d->d_spill, mode == HZ_MODE_WALK ? nullptr : hitem, true, st, err);
        if (rc) return fail(HZ_EDEVICE, err);
        RCHK(hipEventRecord(d->ev[3], st));
    } else {
        // Three classes of cells in the trace order (d_ranges: {first, count}
        // of each): cells next to mesh holes, cells with a mesh-boundary point
        // in reach, the rest.  Three launches run side by side:
        // * border cells (the longest walks: void sectors, descents to -85
        //   deg) run the reference walk by teams of lanes that cast its next
        //   steps together (binary trees; with four-wide trees the
        //   persistent walk runs them first, on the main stream);
        // * the far cells run the lattice search (no boundary point in reach:
        //   their hits are monotone in elevation, so it returns the walk's
        //   values), on a second stream;
        // * the other near cells run the reference walk (persistent kernel)
        //   on the main stream.
        // Side by side, the walk's latency-bound tail and the border teams
        // overlap the lattice items: C2 learned 21.5 ms against 31.6 ms for
        // the three in sequence.
        multi_state = true;
        if (!d->d_state2) RCHK(hipMalloc(&d->d_state2, sizeof(TraceState)));
        if (!d->d_state3) RCHK(hipMalloc(&d->d_state3, sizeof(TraceState)));
        if (d->spill_x_doubles < d->spill_doubles) {  // the far items' spill area
            if (d->d_spill_x) RCHK(hipFree(d->d_spill_x));
            d->d_spill_x = nullptr;
            d->spill_x_doubles = 0;
            RCHK(hipMalloc(&d->d_spill_x, d->spill_doubles * sizeof(double)));
            d->spill_x_doubles = d->spill_doubles;
        }
        const int nb = d->cert.npts;  // cells next to mesh holes
        const bool teams = width == 2;
        constexpr int kBorderTeam = 64, kTeamBlocks = 2048;
        launch_clear_state(d->d_state2, st);
        WalkParams pb = p, pn = p, pf = p;
        pb.range = d->d_ranges + 2;
        pn.range = d->d_ranges + 4;
        pf.range = d->d_ranges + 6;
        if (!teams && nb > 0) {
            rc = launch_trace(d->bvh, d->d_cells, order, pb, HZ_MODE_WALK, 0, block, refill,
                              o.collect_counters != 0, d->d_horizon, d->d_skyview, d->d_state2,
                              d->d_spill, nullptr, true, st, err);
            if (rc) return fail(HZ_EDEVICE, err);
        }
        RCHK(hipEventRecord(d->bev[0], st));
        if (teams) {
            RCHK(hipStreamWaitEvent(d->ystream, d->bev[0], 0));
            if (nb > 0) {
                rc = launch_walk_team(d->bvh, d->d_cells, order, pb, kBorderTeam,
                                      o.collect_counters != 0, d->d_horizon, d->d_skyview,
                                      d->d_state2, nb, kTeamBlocks, d->ystream, err);
                if (rc) return fail(HZ_EDEVICE, err);
            }
            RCHK(hipEventRecord(d->bev[1], d->ystream));
        }
        RCHK(hipStreamWaitEvent(d->xstream, d->bev[0], 0));
        rc = launch_trace(d->bvh, d->d_cells, order, pf, mode, 0, block, refill,
                          o.collect_counters != 0, d->d_horizon, d->d_skyview, d->d_state3,
                          d->d_spill_x, hitem, true, d->xstream, err);
        if (rc) return fail(HZ_EDEVICE, err);
        RCHK(hipEventRecord(d->cev[1], d->xstream));
        rc = launch_trace(d->bvh, d->d_cells, order, pn, HZ_MODE_WALK, 0, block, refill,
                          o.collect_counters != 0, d->d_horizon, d->d_skyview, d->d_state,
                          d->d_spill, nullptr, true, st, err);
        if (rc) return fail(HZ_EDEVICE, err);
        RCHK(hipEventRecord(d->cev[0], st));
        RCHK(hipStreamWaitEvent(st, d->cev[1], 0));
        if (teams) RCHK(hipStreamWaitEvent(st, d->bev[1], 0));
        RCHK(hipEventRecord(d->ev[3], st));
    }
    TraceState ts_host, ts_border, ts_far;
    TraceCounters& c = ts_host.cnt;
    int ranges_host[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    RCHK(hipMemcpyAsync(&ts_host, d->d_state, offsetof(TraceState, pad), hipMemcpyDeviceToHost,
                        st));
    if (multi_state) {
        RCHK(hipMemcpyAsync(&ts_border, d->d_state2, offsetof(TraceState, pad),
                            hipMemcpyDeviceToHost, st));
        RCHK(hipMemcpyAsync(&ts_far, d->d_state3, offsetof(TraceState, pad),
                            hipMemcpyDeviceToHost, st));
        RCHK(hipMemcpyAsync(ranges_host, d->d_ranges, sizeof(ranges_host), hipMemcpyDeviceToHost,
                            st));
    }
    unsigned nodes4 = 0;
    if (d->bvh.d_num_nodes)
        RCHK(hipMemcpyAsync(&nodes4, d->bvh.d_num_nodes, sizeof(nodes4), hipMemcpyDeviceToHost,
                            st));
    RCHK(hipStreamSynchronize(st));
    if (d->bvh.d_num_nodes) d->bvh.num_nodes = (int)nodes4;
    if (multi_state) {  // the border walk's and the far items' rays, visits and flags
        for (const TraceState* o2 : {&ts_border, &ts_far}) {
            c.rays += o2->cnt.rays;
            c.nodes += o2->cnt.nodes;
            c.tris += o2->cnt.tris;
            c.error |= o2->cnt.error;
            c.cyc_refill += o2->cnt.cyc_refill;
            c.cyc_trav += o2->cnt.cyc_trav;
            c.trav_iters += o2->cnt.trav_iters;
            c.spec_rays += o2->cnt.spec_rays;
        }
    }
    float ms_bvh = 0, ms_ord = 0, ms_trace = 0;
    RCHK(hipEventElapsedTime(&ms_bvh, d->ev[0], d->ev[1]));
    if (use_pre) ms_bvh += d->bvh_pre_ms;
    RCHK(hipEventElapsedTime(&ms_ord, d->ev[1], d->ev[2]));
    float ms_pred = 0;  // the predictor stage on its own stream
    if (learned) RCHK(hipEventElapsedTime(&ms_pred, d->pev[1], d->pev[2]));
    RCHK(hipEventElapsedTime(&ms_trace, d->ev[4], d->ev[3]));
    if (cert_run) {
        if (width == 2) RCHK(hipEventElapsedTime(&ms_bwalk, d->bev[0], d->bev[1]));
        RCHK(hipEventElapsedTime(&ms_cwalk, d->bev[0], d->cev[0]));
        RCHK(hipEventElapsedTime(&ms_lattice, d->bev[0], d->cev[1]));
        RCHK(hipEventElapsedTime(&ms_cert, d->ev[5], d->ev[6]));
    }
    double t1 = now_ms();
    d->last_nhori = num_hori;
    d->ran = c.error == 0;
    if (verbose) {
        std::fprintf(stderr, "Building bounding volume hierarchy (BVH): %.2f ms\n", ms_bvh);
        std::fprintf(stderr, "Ray tracing: %.2f ms\n", ms_trace);
        std::fprintf(stderr, "Number of rays shot: %llu\n", (unsigned long long)c.rays);
        std::fprintf(stderr, "Average number of rays per cell and azimuth sector: %.2f\n",
                     (double)c.rays / ((double)d->ncell * num_hori));
    }
    if (stats) {
        stats->num_rays = (long long)c.rays;
        stats->node_visits = (long long)c.nodes;
        stats->tri_tests = (long long)c.tris;
        stats->num_triangles = d->ntri;
        stats->num_mesh_vertices = d->nmeshvert;
        stats->bvh_nodes = d->bvh.num_nodes;
        stats->bvh_depth = d->bvh.depth;
        stats->kernel_error = (int)c.error;
        stats->t_bvh_ms = ms_bvh;
        stats->t_order_ms = ms_ord;
        stats->t_trace_ms = ms_trace;
        stats->t_predict_ms = ms_pred;
        const double cyc = (double)c.cyc_refill + (double)c.cyc_trav;
        stats->refill_frac = cyc > 0 ? (double)c.cyc_refill / cyc : 0.0;
        stats->lanes_per_iter = c.trav_iters ? (double)c.nodes / (double)c.trav_iters : 0.0;
        // cells that ran the walk: the border cells and the other near cells
        stats->cert_walk_cells = cert_run ? (double)ranges_host[3] + (double)ranges_host[5] : 0.0;
        stats->t_certify_ms = ms_cert;
        stats->t_cert_walk_ms = ms_cwalk;
        stats->t_border_walk_ms = ms_bwalk;
        stats->t_cert_lattice_ms = ms_lattice;
        stats->spec_rays = (double)c.spec_rays;
        stats->t_total_ms = t1 - t0;
        stats->lon0 = d->lon0;
        stats->lat0 = d->lat0;
    }
    if (c.error) return fail(HZ_EKERNEL, "trace kernel exceeded its traversal stack or walk limit");
    return HZ_OK;
}

int hz_domain_bvh(hz_domain* d, const hz_options* opt, void* nodes, long long* num_nodes,
                  float* tris, long long* num_tris) {
    g_err.clear();
    if (!d) return fail(HZ_EINVAL, "domain is NULL");
    hz_options o;
    if (opt)
        o = *opt;
    else
        hz_default_options(&o);
    const int leaf = o.leaf_size ? o.leaf_size : 1;
    if (leaf < 1 || leaf > kMaxLeafSize) return fail(HZ_EINVAL, "leaf_size must be in [1, 8]");
    const int width = o.bvh_width ? o.bvh_width : 2;
    if (width != 2 && width != 4) return fail(HZ_EINVAL, "bvh_width must be 2 or 4");
    if (o.bvh_morton < 0 || o.bvh_morton > 1) return fail(HZ_EINVAL, "bvh_morton must be 0 or 1");
    RCHK(hipSetDevice(d->device));
    std::string err;
    d->bvh.nodes = d->d_nodes;
    d->bvh.tris = d->d_tris;
    if (o.bvh_sah && width != 2) return fail(HZ_EINVAL, "bvh_sah needs the binary tree (bvh_width 2)");
    int rc = lbvh_build(d->work, d->d_vert, d->d_tri, (int)d->ntri, leaf, o.bvh_morton, width,
                        d->stream, d->bvh, false, err, o.bvh_sah == 1);
    if (rc) return fail(rc == -3 ? HZ_ENOMEM : (rc == -1 ? HZ_EINVAL : HZ_EDEVICE), err);
    long long count = 0;
    if (d->ntri > 0) {
        if (width == 4) {
            unsigned c = 0;
            RCHK(hipMemcpyAsync(&c, d->bvh.d_num_nodes, sizeof(c), hipMemcpyDeviceToHost, d->stream));
            RCHK(hipStreamSynchronize(d->stream));
            count = c;
        } else {
            count = std::max<long long>(1, d->ntri - 1);
        }
    }
    if (num_nodes) *num_nodes = count;
    if (num_tris) *num_tris = d->ntri;
    if (nodes && count)
        RCHK(hipMemcpyAsync(nodes, d->d_nodes, (size_t)count * sizeof(BvhNode),
                            hipMemcpyDeviceToHost, d->stream));
    if (tris && d->ntri)
        RCHK(hipMemcpyAsync(tris, d->d_tris, (size_t)d->ntri * 3 * sizeof(float4),
                            hipMemcpyDeviceToHost, d->stream));
    RCHK(hipStreamSynchronize(d->stream));
    return HZ_OK;
}

int hz_domain_fetch(hz_domain* d, double* horizon, double* skyview, hz_stats* stats) {
    g_err.clear();
    if (!d) return fail(HZ_EINVAL, "domain is NULL");
    if (!d->ran) return fail(HZ_EINVAL, "no successful hz_domain_run to fetch");
    RCHK(hipSetDevice(d->device));
    double t0 = now_ms();
    if (horizon)
        RCHK(d->stager.copy(horizon, d->d_horizon,
                            (size_t)d->last_nhori * d->ncell * sizeof(double), false, d->stream));
    if (skyview)
        RCHK(d->stager.copy(skyview, d->d_skyview, (size_t)d->ncell * sizeof(double), false,
                            d->stream));
    if (stats) stats->t_download_ms = now_ms() - t0;
    return HZ_OK;
}

// Device-buffer cache of the one-shot calls (horizon_svf_comp and its
// variants, option cache = 1): per device, the domain of the last call is
// kept, buffers, streams, events and pinned staging included, and the next
// call whose host prologue gives the same sizes (cells, triangles, mesh
// vertices) loads its scene into them instead of allocating.  The kept
// domain is taken out of the cache while a call uses it, so concurrent calls
// never share one.  hz_cache_release frees it.
namespace {
constexpr int kCacheDevices = 64;
std::mutex g_cache_mu;
hz_domain* g_cache[kCacheDevices] = {};

// The kept domain of device dev if its sizes match (removed from the
// cache); a kept domain of other sizes is freed first, so that the call's
// own allocation does not have to fit beside it.
hz_domain* cache_take(int dev, int ncell, long long ntri, long long nmv) {
    if (dev < 0 || dev >= kCacheDevices) return nullptr;
    hz_domain* d = nullptr;
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        d = g_cache[dev];
        g_cache[dev] = nullptr;
    }
    if (d && (d->ncell != ncell || d->ntri != ntri || d->nmeshvert != nmv || !d->host_scene)) {
        hz_domain_destroy(d);
        d = nullptr;
    }
    return d;
}

// The host scene arrays of the last one-shot call (about 150 bytes per
// cell), kept so that the next prologue writes into memory that is already
// mapped instead of faulting in fresh pages.
std::mutex g_host_mu;
HostScene* g_host_kept = nullptr;

HostScene* host_take() {
    HostScene* h = nullptr;
    {
        std::lock_guard<std::mutex> g(g_host_mu);
        h = g_host_kept;
        g_host_kept = nullptr;
    }
    return h ? h : new HostScene();
}

void host_put(HostScene* h) {
    {
        std::lock_guard<std::mutex> g(g_host_mu);
        if (!g_host_kept) {
            g_host_kept = h;
            h = nullptr;
        }
    }
    delete h;
}

// Keep d for the next call on its device (a domain kept meanwhile by a
// concurrent call is freed).
void cache_put(hz_domain* d) {
    if (!d || !d->host_scene || d->device < 0 || d->device >= kCacheDevices) {
        hz_domain_destroy(d);
        return;
    }
    hz_domain* old = nullptr;
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        old = g_cache[d->device];
        g_cache[d->device] = d;
    }
    hz_domain_destroy(old);
}

// Load a host scene of the same sizes into a kept domain: upload the FP32
// scene and start the default tree build on the device, then upload the FP64
// ENU cell points and build the cell frames from them on the device (the
// host's frame arithmetic, the same bits), with the hole flags from the
// host's list (hole_cells).  What was
// derived from the previous scene (certificate index, slope adjacency,
// predictor, cell flags, run state) is dropped.
int domain_reload(hz_domain* d, HostScene& hs, const double* hsurf, const int* cov,
                  int num_cell, int num_vertex, int grid_type, double ray_org_elev,
                  hz_stats* stats, double t0, double t1) {
    hipError_t e = hipSetDevice(d->device);
    d->ran = false;
    d->last_nhori = 0;
    d->cert.dist_key = -1.0f;
    d->near_flags = false;
    d->predictor = nullptr;
    d->have_pred = false;
    d->bvh_pre = false;
    if (e == hipSuccess && d->d_adj_off) e = hipFree(d->d_adj_off);
    if (e == hipSuccess && d->d_adj) e = hipFree(d->d_adj);
    d->d_adj_off = nullptr;
    d->d_adj = nullptr;
    d->grid_type = grid_type;
    d->hmin = d->hmax = hsurf[0];
    for (int i = 1; i < num_cell; i++) {
        d->hmin = std::min(d->hmin, hsurf[i]);
        d->hmax = std::max(d->hmax, hsurf[i]);
    }
    d->lon0 = hs.lon0;
    d->lat0 = hs.lat0;
    std::memcpy(d->bounds_xy, hs.bounds_xy, sizeof(d->bounds_xy));
    const size_t nv = hs.vert32.size(), nt = hs.tri.size();
    for (auto& x : d->pre_ev)
        if (e == hipSuccess && !x) e = hipEventCreate(&x);
    if (e == hipSuccess && nv)
        e = d->stager.copy(d->d_vert, hs.vert32.data(), nv * sizeof(float), true, d->stream);
    if (e == hipSuccess && nt)
        e = d->stager.copy(d->d_tri, hs.tri.data(), nt * sizeof(int), true, d->stream);
    std::string err;
    if (e == hipSuccess) e = hipEventRecord(d->pre_ev[0], d->stream);
    if (e == hipSuccess) {
        d->bvh.nodes = d->d_nodes;
        d->bvh.tris = d->d_tris;
        const int rc = lbvh_build(d->work, d->d_vert, d->d_tri, (int)d->ntri, 1, 0, 2, d->stream,
                                  d->bvh, false, err);
        if (rc) return fail(rc == -3 ? HZ_ENOMEM : HZ_EDEVICE, err);
        e = hipEventRecord(d->pre_ev[1], d->stream);
    }
    // the cell frames: built on the device from the FP64 ENU cell points (the
    // host's frame arithmetic, same bits), then the hole flags from the host's
    // list of cells next to mesh holes
    if (e == hipSuccess && !d->d_pcell)
        e = hipMalloc(&d->d_pcell, (size_t)num_cell * 3 * sizeof(double));
    if (e == hipSuccess)
        e = d->stager.copy(d->d_pcell, hs.work.data(), (size_t)num_cell * 3 * sizeof(double), true,
                           d->stream);
    if (e == hipSuccess)
        launch_frames_from_enu(d->d_pcell, num_cell, hs.centre, hs.pole, ray_org_elev, d->d_cells,
                               d->stream);
    const double tf0 = now_ms();
    hole_cells(cov, num_cell, num_vertex, hs.holes);
    const double tf1 = now_ms();
    const int nh = (int)hs.holes.size();
    if (e == hipSuccess && (size_t)nh > d->holes_cap) {
        if (d->d_holes) e = hipFree(d->d_holes);
        d->d_holes = nullptr;
        d->holes_cap = 0;
        if (e == hipSuccess) e = hipMalloc(&d->d_holes, (size_t)nh * sizeof(int));
        if (e == hipSuccess) d->holes_cap = (size_t)nh;
    }
    if (e == hipSuccess && nh > 0)
        e = hipMemcpyAsync(d->d_holes, hs.holes.data(), (size_t)nh * sizeof(int),
                           hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) {
        launch_flag_cells(d->d_holes, nh, d->d_cells, d->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&d->bvh_pre_ms, d->pre_ev[0], d->pre_ev[1]);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? HZ_ENOMEM : HZ_EDEVICE,
                    std::string("reloading a kept domain: ") + hipGetErrorString(e));
    d->bvh_pre = true;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->num_triangles = d->ntri;
        stats->num_mesh_vertices = d->nmeshvert;
        stats->t_prologue_ms = (t1 - t0) + (tf1 - tf0);
        stats->t_upload_ms = (now_ms() - t1) - (tf1 - tf0);
        stats->lon0 = d->lon0;
        stats->lat0 = d->lat0;
    }
    return HZ_OK;
}
}  // namespace

void hz_cache_release(void) {
    hz_domain* kept[kCacheDevices];
    {
        std::lock_guard<std::mutex> g(g_cache_mu);
        for (int i = 0; i < kCacheDevices; i++) {
            kept[i] = g_cache[i];
            g_cache[i] = nullptr;
        }
    }
    for (hz_domain* d : kept) hz_domain_destroy(d);
    HostScene* h = nullptr;
    {
        std::lock_guard<std::mutex> g(g_host_mu);
        h = g_host_kept;
        g_host_kept = nullptr;
    }
    delete h;
}

int horizon_svf_comp_ex(const double* clon, const double* clat, const double* hsurf,
                        const double* vlon, const double* vlat, const int* cov, double* horizon,
                        double* skyview, int num_cell, int num_vertex, int num_hori,
                        int grid_type, double dist_search, double ray_org_elev,
                        int refine_factor, int svf_type, const hz_options* opt,
                        hz_stats* stats) {
    double t0 = now_ms();
    g_err.clear();
    if (!horizon || !skyview) return fail(HZ_EINVAL, "NULL output array");
    int rc = check_run_args(num_hori, dist_search, refine_factor, svf_type);
    if (rc) {
        fill_nan(horizon, skyview, num_hori, num_cell);
        return rc;
    }
    hz_options o;
    if (opt)
        o = *opt;
    else
        hz_default_options(&o);
    bool verbose = o.verbose || verbose_env();
    if (verbose) {
        std::fprintf(stderr, "-------------------------------------------------------------------------------\n");
        std::fprintf(stderr, "Horizon and SVF computation with HIP on gfx950\n");
        std::fprintf(stderr, "-------------------------------------------------------------------------------\n");
    }
    hz_stats local;
    hz_stats* s = stats ? stats : &local;
    std::memset(s, 0, sizeof(*s));
    hz_domain* d = nullptr;
    if (o.cache != 0 && o.cache != 1) rc = fail(HZ_EINVAL, "cache must be 0 or 1");
    if (!rc && o.cache == 1 && o.prologue == HZ_PROLOGUE_HOST) {
        // host prologue, then the kept domain of these sizes if there is one
        rc = clon && clat && hsurf && (num_vertex <= 0 || (vlon && vlat && cov))
                 ? check_grid_args(cov, num_cell, num_vertex, grid_type, ray_org_elev)
                 : fail(HZ_EINVAL, "NULL input array");
        HostScene* hsp = host_take();
        HostScene& hs = *hsp;
        std::string err;
        const double tp0 = now_ms();
        if (!rc) {
            rc = prologue_scene(clon, clat, hsurf, vlon, vlat, cov, num_cell, num_vertex,
                                grid_type, hs, err);
            if (rc) rc = fail(HZ_EINVAL, err);
        }
        if (!rc) {
            const double tp1 = now_ms();
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess) dev = -1;
            d = cache_take(dev, num_cell, (long long)hs.tri.size() / 3, hs.num_mesh_vertices);
            if (d) {
                rc = domain_reload(d, hs, hsurf, cov, num_cell, num_vertex, grid_type,
                                   ray_org_elev, s, tp0, tp1);
                if (rc) {
                    hz_domain_destroy(d);
                    d = nullptr;
                }
            } else {
                prologue_frames(cov, num_cell, num_vertex, ray_org_elev, hs);
                rc = make_domain(clon, clat, hsurf, vlon, vlat, cov, num_cell, num_vertex,
                                 grid_type, ray_org_elev, HZ_PROLOGUE_HOST, hs, tp0, &d, s);
            }
        }
        host_put(hsp);
    } else if (!rc) {
        rc = hz_domain_create(clon, clat, hsurf, vlon, vlat, cov, num_cell, num_vertex,
                              grid_type, ray_org_elev, &o, &d, s);
    }
    if (rc) {
        fill_nan(horizon, skyview, num_hori, num_cell);
        return rc;
    }
    double tp = s->t_prologue_ms, tu = s->t_upload_ms;
    if (verbose) {
        std::fprintf(stderr, "Number of triangles in mesh: %lld\n", d->ntri);
        std::fprintf(stderr, "Origin of ENU coordinate system: %.4f deg lat, %.4f deg lon\n",
                     d->lat0 * 180.0 / 3.14159265358979323846, d->lon0 * 180.0 / 3.14159265358979323846);
    }
    rc = hz_domain_run(d, num_hori, dist_search, refine_factor, svf_type, &o, s);
    if (rc == 0) rc = hz_domain_fetch(d, horizon, skyview, s);
    std::string keep = g_err;
    if (rc == 0 && o.cache == 1)
        cache_put(d);
    else
        hz_domain_destroy(d);
    g_err = keep;
    s->t_prologue_ms = tp;
    s->t_upload_ms = tu;
    s->t_total_ms = now_ms() - t0;
    if (rc) {
        fill_nan(horizon, skyview, num_hori, num_cell);
        return rc;
    }
    return HZ_OK;
}

void horizon_svf_comp(double* clon, double* clat, double* hsurf, double* vlon, double* vlat,
                      int* cells_of_vertex, double* horizon, double* skyview, int num_cell,
                      int num_vertex, int num_hori, int grid_type, double dist_search,
                      double ray_org_elev, int refine_factor, int svf_type) {
    int rc = horizon_svf_comp_ex(clon, clat, hsurf, vlon, vlat, cells_of_vertex, horizon, skyview,
                                 num_cell, num_vertex, num_hori, grid_type, dist_search,
                                 ray_org_elev, refine_factor, svf_type, nullptr, nullptr);
    if (rc) std::fprintf(stderr, "horizon_svf_comp failed (%d): %s\n", rc, g_err.c_str());
}

void horizon_svf_comp_fortran(double* clon, double* clat, double* hsurf, double* vlon,
                              double* vlat, int* cells_of_vertex, double* horizon,
                              double* skyview, int num_cell, int num_vertex, int num_hori,
                              int grid_type, double dist_search, double ray_org_elev,
                              int refine_factor, int svf_type, char* buffer, int* buffer_len) {
    hz_stats st;
    std::memset(&st, 0, sizeof(st));
    int rc = horizon_svf_comp_ex(clon, clat, hsurf, vlon, vlat, cells_of_vertex, horizon, skyview,
                                 num_cell, num_vertex, num_hori, grid_type, dist_search,
                                 ray_org_elev, refine_factor, svf_type, nullptr, &st);
    const std::string text = run_report(st, num_cell, num_hori, grid_type, svf_type, rc);
    if (buffer_len) {
        const int n = buffer && *buffer_len > 0 ? std::min(*buffer_len, (int)text.size()) : 0;
        if (n > 0) std::memcpy(buffer, text.data(), (size_t)n);
        *buffer_len = n;
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched domains
// ---------------------------------------------------------------------------

struct hz_batch {
    std::vector<hz_domain*> doms;
    int device = 0;
    hipStream_t st = nullptr;
    BatchDomain* d_table = nullptr;
    TraceState* d_state = nullptr;
    double* d_spill = nullptr;
    size_t spill_cap = 0;
    int total = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // configuration the graph was captured for
    int g_mode = -1, g_block = 0, g_refill = 0, g_chunk = 0, g_multi = -1, g_width = 0,
        g_nhori = 0, g_rf = 0, g_svf = 0;
    double g_dist = 0;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
};

extern "C" {

void hz_batch_destroy(hz_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->st) (void)hipStreamSynchronize(b->st);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* p[] = {b->d_table, b->d_state, b->d_spill};
    for (void* q : p)
        if (q) (void)hipFree(q);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

#define BCHK(call)                                                                     \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            hz_batch_destroy(b);                                                       \
            return fail(e_ == hipErrorOutOfMemory ? HZ_ENOMEM : HZ_EDEVICE,           \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
        }                                                                              \
    } while (0)

int hz_batch_create(hz_domain* const* doms, int n, hz_batch** out) {
    g_err.clear();
    if (!out || !doms || n < 1) return fail(HZ_EINVAL, "need at least one domain");
    *out = nullptr;
    hz_batch* b = new hz_batch();
    long long total = 0;
    for (int i = 0; i < n; i++) {
        if (!doms[i]) {
            delete b;
            return fail(HZ_EINVAL, "NULL domain in batch");
        }
        if (i && doms[i]->device != doms[0]->device) {
            delete b;
            return fail(HZ_EINVAL, "all domains of a batch must live on one device");
        }
        b->doms.push_back(doms[i]);
        total += doms[i]->ncell;
    }
    if (total > (long long)0x7fffffff - 4096) {
        delete b;
        return fail(HZ_EINVAL, "batch too large");
    }
    b->total = (int)total;
    b->device = doms[0]->device;
    BCHK(hipSetDevice(b->device));
    BCHK(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    for (auto& e : b->ev) BCHK(hipEventCreate(&e));
    BCHK(hipMalloc(&b->d_table, n * sizeof(BatchDomain)));
    BCHK(hipMalloc(&b->d_state, sizeof(TraceState)));
    *out = b;
    return HZ_OK;
}

#undef BCHK

// One pass over all domains of the batch: per-domain LBVH build and cell
// ordering (stream-ordered), then ONE persistent trace launch over every
// cell -- replayed from a captured hipGraph when use_graph (captured on the
// first call of a configuration).  Outputs stay in each domain's buffers;
// hz_domain_fetch reads them.
int hz_batch_run(hz_batch* b, int num_hori, double dist_search, int refine_factor, int svf_type,
                 const hz_options* opt, int use_graph, hz_stats* stats) {
    g_err.clear();
    if (!b) return fail(HZ_EINVAL, "batch is NULL");
    int rc = check_run_args(num_hori, dist_search, refine_factor, svf_type);
    if (rc) return rc;
    hz_options o;
    if (opt)
        o = *opt;
    else
        hz_default_options(&o);
    if (o.search_mode != HZ_MODE_WALK && o.search_mode != HZ_MODE_CLIMB)
        return fail(HZ_EINVAL, "batches run the walk or the climb search");
    if (o.search_mode == HZ_MODE_CLIMB && o.certify != 0)
        return fail(HZ_EINVAL,
                    "batches run the plain climb only (whole cells, no certificate): pass certify 0");
    int block = o.block_size ? o.block_size : 128;
    if (block != 64 && block != 128 && block != 256)
        return fail(HZ_EINVAL, "block_size must be 64, 128 or 256");
    int leaf = o.leaf_size ? o.leaf_size : 1;
    if (leaf < 1 || leaf > kMaxLeafSize) return fail(HZ_EINVAL, "leaf_size must be in [1, 8]");
    int refill = o.refill ? o.refill : kDefaultRefill;
    if (refill < 1 || refill > 64) return fail(HZ_EINVAL, "refill must be in [1, 64]");
    int chunk = o.queue_chunk ? o.queue_chunk : 64;  // cells per queue fetch
    if (chunk < 1 || chunk > 4096) return fail(HZ_EINVAL, "queue_chunk must be in [1, 4096]");
    int width = o.bvh_width ? o.bvh_width : 2;
    if (width != 2 && width != 4) return fail(HZ_EINVAL, "bvh_width must be 2 or 4");
    RCHK(hipSetDevice(b->device));
    hipStream_t st = b->st;
    std::string err;
    double t0 = now_ms();
    RCHK(hipEventRecord(b->ev[0], st));
    std::vector<BatchDomain> table(b->doms.size());
    int start = 0;
    for (size_t i = 0; i < b->doms.size(); i++) {
        hz_domain* d = b->doms[i];
        if (num_hori > d->out_nhori) {
            if (d->d_horizon) RCHK(hipFree(d->d_horizon));
            d->d_horizon = nullptr;
            RCHK(hipMalloc(&d->d_horizon, (size_t)num_hori * d->ncell * sizeof(double)));
            d->out_nhori = num_hori;
        }
        d->bvh.nodes = d->d_nodes;
        d->bvh.tris = d->d_tris;
        rc = lbvh_build(d->work, d->d_vert, d->d_tri, (int)d->ntri, leaf, o.bvh_morton, width, st,
                        d->bvh, false, err, o.bvh_sah == 1 && width == 2);
        if (rc) return fail(rc == -3 ? HZ_ENOMEM : (rc == -1 ? HZ_EINVAL : HZ_EDEVICE), err);
        if (o.cell_order) {
            rc = order_cells(d->work, d->d_cells, d->ncell, d->d_order, st, err);
            if (rc) return fail(HZ_EDEVICE, err);
        }
        BatchDomain& e = table[i];
        e.nodes = d->d_nodes;
        e.tris = d->d_tris;
        e.cells = d->d_cells;
        e.order = o.cell_order ? d->d_order : nullptr;
        e.horizon = d->d_horizon;
        e.skyview = d->d_skyview;
        e.num_tri = (int)d->ntri;
        e.ncell = d->ncell;
        e.start = start;
        e.pad = 0;
        start += d->ncell;
    }
    RCHK(hipMemcpyAsync(b->d_table, table.data(), table.size() * sizeof(BatchDomain),
                        hipMemcpyHostToDevice, st));
    const bool multi = leaf > 1;
    const int mode = o.search_mode;
    size_t need = (size_t)batch_grid(mode, block, multi, width, b->total) * S_FIELDS_HOST;
    if (need > b->spill_cap) {
        if (b->d_spill) RCHK(hipFree(b->d_spill));
        b->d_spill = nullptr;
        RCHK(hipMalloc(&b->d_spill, need * sizeof(double)));
        b->spill_cap = need;
        if (b->exec) {  // the captured graph points at the old spill area
            (void)hipGraphExecDestroy(b->exec);
            (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
        }
    }
    WalkParams p = make_walk_params(num_hori, refine_factor, svf_type, dist_search, 0);
    p.elev_cs = b->doms[0]->d_elev_cs;
    p.sin_h = reinterpret_cast<const double*>(b->doms[0]->d_elev_cs + (kMaxLattice + 1));
    RCHK(hipEventRecord(b->ev[1], st));
    if (use_graph) {
        const bool same = b->exec && b->g_mode == mode && b->g_block == block &&
                          b->g_refill == refill && b->g_chunk == chunk && b->g_multi == (int)multi &&
                          b->g_width == width &&
                          b->g_nhori == num_hori && b->g_rf == refine_factor &&
                          b->g_svf == svf_type && b->g_dist == dist_search;
        if (!same) {
            if (b->exec) {
                (void)hipGraphExecDestroy(b->exec);
                (void)hipGraphDestroy(b->graph);
                b->exec = nullptr;
                b->graph = nullptr;
            }
            RCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = launch_trace_batch(b->d_table, (int)b->doms.size(), b->total, multi, width, p, mode, block,
                                    refill, chunk, b->d_state, b->d_spill, st, err);
            hipGraph_t g = nullptr;
            hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc || ec != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(HZ_EDEVICE, rc ? err : std::string("hipStreamEndCapture failed"));
            }
            b->graph = g;
            RCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->g_mode = mode;
            b->g_block = block;
            b->g_refill = refill;
            b->g_chunk = chunk;
            b->g_multi = multi;
            b->g_width = width;
            b->g_nhori = num_hori;
            b->g_rf = refine_factor;
            b->g_svf = svf_type;
            b->g_dist = dist_search;
        }
        RCHK(hipGraphLaunch(b->exec, st));
    } else {
        rc = launch_trace_batch(b->d_table, (int)b->doms.size(), b->total, multi, width, p, mode, block,
                                refill, chunk, b->d_state, b->d_spill, st, err);
        if (rc) return fail(HZ_EDEVICE, err);
    }
    RCHK(hipEventRecord(b->ev[2], st));
    TraceCounters c;
    RCHK(hipMemcpyAsync(&c, &b->d_state->cnt, sizeof(c), hipMemcpyDeviceToHost, st));
    RCHK(hipStreamSynchronize(st));
    float ms_build = 0, ms_trace = 0;
    RCHK(hipEventElapsedTime(&ms_build, b->ev[0], b->ev[1]));
    RCHK(hipEventElapsedTime(&ms_trace, b->ev[1], b->ev[2]));
    for (hz_domain* d : b->doms) {
        d->last_nhori = num_hori;
        d->ran = c.error == 0;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->num_rays = (long long)c.rays;
        stats->kernel_error = (int)c.error;
        stats->t_bvh_ms = ms_build;
        stats->t_trace_ms = ms_trace;
        stats->t_total_ms = now_ms() - t0;
    }
    if (c.error) return fail(HZ_EKERNEL, "trace kernel exceeded its traversal stack or walk limit");
    return HZ_OK;
}

}  // extern "C"
