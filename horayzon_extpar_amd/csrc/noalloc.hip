// The reference walk with no device allocation: hz_walk_noalloc.
//
// The same device pass as hz_domain_run with search_mode 0 (LBVH build,
// Morton cell order, the persistent walk kernel with its fused sector mean
// and SVF), over a scene the caller holds in device memory and into output
// buffers the caller owns.  Every intermediate buffer (the tree and its build
// work, the packed cell frames, the cell order, the walk's spill area and
// queue state, the lattice table) is carved out of one file-scope device
// array, `scratch`, so that a call makes no hipMalloc, hipFree or other
// allocating call: it enqueues kernels on the caller's stream and returns.
//
// The array is a host-side stand-in in host code: its device address comes
// from hipGetSymbolAddress, once per device, at the first call on that device
// (which also writes the constant lattice table with a synchronous copy), and
// only that address reaches kernels.  Each device holds its own copy of the
// array until the process ends.  All calls on one device share it: calls
// that overlap on one device (other streams, other host threads) race.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/horayzon.h"
#include "hz_internal.h"

using namespace hz;

namespace {

constexpr long long kMaxTri = HZ_WALK_NOALLOC_MAX_TRIANGLES;
constexpr int kMaxCell = HZ_WALK_NOALLOC_MAX_CELLS;
constexpr size_t kScratchBytes = (size_t)HZ_WALK_NOALLOC_SCRATCH_MIB << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Byte offsets of the buffers in one device's array (256-byte aligned).
struct Layout {
    size_t table, state, cells, order, nodes, tris, spill, work, total;
};

struct DeviceScratch {
    unsigned char* base = nullptr;  // hipGetSymbolAddress(scratch) on this device
    Layout at{};
    LbvhWork work;  // borrowed: the build buffers inside the array
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

__global__ void k_pack_frames(const double* __restrict__ n, const double* __restrict__ north,
                              const float* __restrict__ org, int ncell, CellFrame* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncell) return;
    CellFrame f;
    for (int a = 0; a < 3; a++) {
        f.n[a] = n[3 * (size_t)i + a];
        f.north[a] = north[3 * (size_t)i + a];
        f.org[a] = org[3 * (size_t)i + a];
    }
    f.pad = 0;  // no mesh-hole flags: only the order of the work changes, not a result
    out[i] = f;
}

// First call on device dev: the array's device address, the layout, and the
// lattice table written into it.  g_mu is held by the caller.
int setup(int dev, std::string& err) {
    DeviceScratch& d = g_dev[dev];
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return HZ_EDEVICE;
    }
    Layout L;
    size_t off = 0;
    L.table = off;
    off += up256((kMaxLattice + 1) * (sizeof(double2) + sizeof(double)));
    L.state = off;
    off += up256(sizeof(TraceState));
    L.cells = off;
    off += up256((size_t)kMaxCell * sizeof(CellFrame));
    L.order = off;
    off += up256((size_t)kMaxCell * sizeof(int));
    L.nodes = off;
    off += up256((size_t)(kMaxTri - 1) * sizeof(BvhNode));
    L.tris = off;
    off += up256((size_t)kMaxTri * 3 * sizeof(float4));
    L.spill = off;
    off += up256(trace_spill_doubles(kMaxCell) * sizeof(double));
    L.work = off;
    LbvhWork w;
    const int wn = (int)std::max<long long>(kMaxTri, kMaxCell);
    const size_t wb = lbvh_bind(w, wn, nullptr, err);
    if (!wb) return HZ_EDEVICE;
    off += up256(wb);
    L.total = off;
    if (L.total > kScratchBytes) {
        err = "no-alloc scratch layout needs " + std::to_string(L.total) + " bytes, the array holds " +
              std::to_string(kScratchBytes);
        return HZ_EDEVICE;
    }
    unsigned char* base = static_cast<unsigned char*>(p);
    lbvh_bind(d.work, wn, reinterpret_cast<char*>(base + L.work), err);
    double2 cs[kMaxLattice + 1];
    double sh[kMaxLattice + 1];
    lattice_table(make_walk_params(24, 1, 2, 4e4, 1).hori_acc, cs, sh);
    e = hipMemcpy(base + L.table, cs, sizeof(cs), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(base + L.table + sizeof(cs), sh, sizeof(sh), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        err = std::string("lattice table: ") + hipGetErrorString(e);
        return HZ_EDEVICE;
    }
    d.at = L;
    d.base = base;
    return HZ_OK;
}

}  // namespace

extern "C" int hz_walk_noalloc(const float* vert32, long long num_mesh_vertices, const int* tri,
                               long long num_triangles, const double* frame_n,
                               const double* frame_north, const float* org, int num_cell,
                               int num_hori, double dist_search, int refine_factor,
                               int svf_type, double* horizon, double* skyview, void* stream) {
    api_clear_error();
    int rc = api_check_run_args(num_hori, dist_search, refine_factor, svf_type);
    if (rc) return rc;
    if (num_cell < 1) return api_fail(HZ_EINVAL, "num_cell must be >= 1");
    if (num_triangles < 0 || num_mesh_vertices < 0)
        return api_fail(HZ_EINVAL, "num_triangles and num_mesh_vertices must be >= 0");
    if (num_cell > kMaxCell || num_triangles > kMaxTri)
        return api_fail(HZ_ENOMEM,
                        "hz_walk_noalloc holds at most " + std::to_string(kMaxCell) +
                            " cells and " + std::to_string(kMaxTri) + " triangles (asked " +
                            std::to_string(num_cell) + " and " + std::to_string(num_triangles) +
                            "); use hz_domain_run or horizon_svf_comp");
    if (!frame_n || !frame_north || !org || !horizon || !skyview ||
        (num_triangles > 0 && (!vert32 || !tri || num_mesh_vertices < 3)))
        return api_fail(HZ_EINVAL, "hz_walk_noalloc: NULL array");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return api_fail(HZ_EDEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return api_fail(HZ_EDEVICE, "device ordinal out of range");
    std::string err;
    DeviceScratch* d = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (!g_dev[dev].base) {
            rc = setup(dev, err);
            if (rc) return api_fail(rc, err);
        }
        d = &g_dev[dev];
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned char* b = d->base;
    const Layout& L = d->at;
    CellFrame* cells = reinterpret_cast<CellFrame*>(b + L.cells);
    hipLaunchKernelGGL(k_pack_frames, dim3((num_cell + 255) / 256), dim3(256), 0, st, frame_n,
                       frame_north, org, num_cell, cells);
    DeviceBvh bvh;
    bvh.nodes = reinterpret_cast<BvhNode*>(b + L.nodes);
    bvh.tris = reinterpret_cast<float4*>(b + L.tris);
    LbvhWork& w = d->work;
    rc = lbvh_build(w, vert32, tri, (int)num_triangles, 1, 0, 2, st, bvh, false, err);
    if (rc) return api_fail(rc == -3 ? HZ_ENOMEM : (rc == -1 ? HZ_EINVAL : HZ_EDEVICE), err);
    int* order = reinterpret_cast<int*>(b + L.order);
    rc = order_cells(w, cells, num_cell, order, st, err);
    if (rc) return api_fail(HZ_EDEVICE, err);
    WalkParams p = make_walk_params(num_hori, refine_factor, svf_type, dist_search, num_cell);
    p.elev_cs = reinterpret_cast<const double2*>(b + L.table);
    p.sin_h = reinterpret_cast<const double*>(b + L.table + (kMaxLattice + 1) * sizeof(double2));
    rc = launch_trace(bvh, cells, order, p, HZ_MODE_WALK, 0, 128, 40, false, horizon, skyview,
                      reinterpret_cast<TraceState*>(b + L.state),
                      reinterpret_cast<double*>(b + L.spill), nullptr, true, st, err);
    if (rc) return api_fail(HZ_EDEVICE, err);
    return HZ_OK;
}
