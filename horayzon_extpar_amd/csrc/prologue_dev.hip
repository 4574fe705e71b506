// Device prologue: ICON cells / vertices -> triangle mesh -> spherical ECEF
// -> local ENU -> FP32 scene and per-cell ray frames, on the GPU.
//
// Same steps and operation order as the host prologue (prologue.cpp), which
// follows mo_lradtopo_horayzon.cpp:
//   mesh (triangles per ICON vertex, added vertices)   :228-317
//   lonlat2ecef                                          :158-169
//   ENU origin + ecef2enu                                :178-201, :610-644
//   per-cell frame / ray origin                          :693-722
// Differences from the host path, both in how values are obtained, not in
// the arithmetic that uses them:
//   * sin / cos of the vertex longitudes / latitudes are correctly rounded
//     (crtrig.h); glibc's agree with them except on arguments within ~0.02
//     ulp of a rounding midpoint (about 0.2% of arguments), where the two
//     FP64 values differ in the last bit;
//   * the cells around an ICON vertex are ordered by the device's atan2, so
//     the order can differ from the host's only for angles within a few ulp.
// The ENU origin is the host's (enu_frame_host, glibc sin / cos, sequential
// sum in cell order), computed on host threads while the device builds the
// mesh, so the transform constants are the host prologue's bit for bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include <rocprim/rocprim.hpp>

#include "crtrig.h"
#include "hz_internal.h"

namespace hz {
namespace {

constexpr int kBlockP = 256;
constexpr double kPiP = 3.14159265358979323846;
constexpr double kRadEarthP = 6371229.0;  // reference :584

inline int gridp(long long n) { return (int)((n + kBlockP - 1) / kBlockP); }

// triangles and added mesh vertices of each ICON vertex (reference :243-312);
// the entry at nvert stays 0 so that an exclusive scan ends with the totals
__global__ void k_valence(const int* __restrict__ cov, int nvert, int grid_type,
                          int* __restrict__ ntri_v, int* __restrict__ nadd_v) {
    const int v = blockIdx.x * kBlockP + threadIdx.x;
    if (v > nvert) return;
    if (v == nvert) {
        ntri_v[v] = 0;
        nadd_v[v] = 0;
        return;
    }
    int n = 0;
    for (int j = 0; j < 6; j++)
        if (cov[(size_t)nvert * j + v] != -2) n++;
    if (grid_type == 0) {
        ntri_v[v] = n >= 3 ? n - 2 : 0;
        nadd_v[v] = 0;
    } else {
        ntri_v[v] = n == 6 ? 6 : 0;
        nadd_v[v] = n == 6 ? 1 : 0;
    }
}

// Fills the triangles of each ICON vertex and its added mesh vertex (lon,
// lat, mean elevation of the six cells), with the valid cells ordered
// clockwise from the latitude axis by atan2(dlon, dlat) in [0, 2 pi); equal
// angles keep their slot order (insertion sort, as the host prologue).
__global__ void k_mesh(const double* __restrict__ clon, const double* __restrict__ clat,
                       const double* __restrict__ hsurf, const double* __restrict__ vlon,
                       const double* __restrict__ vlat, const int* __restrict__ cov, int nvert,
                       int ncell, int grid_type, const int* __restrict__ toff,
                       const int* __restrict__ aoff, int* __restrict__ tri,
                       double* __restrict__ P) {
    const int v = blockIdx.x * kBlockP + threadIdx.x;
    if (v >= nvert) return;
    const int nt = toff[v + 1] - toff[v];
    if (nt == 0) return;
    double ang[6];
    int c6[6];
    int n = 0;
    double hs = 0.0;
    for (int j = 0; j < 6; j++) {
        const int c = cov[(size_t)nvert * j + v];
        if (c == -2) continue;
        double a = atan2(clon[c] - vlon[v], clat[c] - vlat[v]);
        if (a < 0.0) a += 2.0 * kPiP;
        ang[n] = a;
        c6[n] = c;
        hs += hsurf[c];
        n++;
    }
    int idx[6];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++) {
        const int t = idx[i];
        int j = i - 1;
        while (j >= 0 && ang[t] < ang[idx[j]]) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = t;
    }
    int cell[6];
    for (int i = 0; i < n; i++) cell[i] = c6[idx[i]];
    int* t = tri + (size_t)toff[v] * 3;
    if (grid_type == 0) {
        for (int j = 1; j + 1 < n; j++) {  // fan about the first cell
            *t++ = cell[0];
            *t++ = cell[j];
            *t++ = cell[j + 1];
        }
    } else {
        const long long add = (long long)ncell + aoff[v];
        P[3 * add] = vlon[v];
        P[3 * add + 1] = vlat[v];
        P[3 * add + 2] = hs / 6.0;
        for (int j = 0; j < 6; j++) {
            *t++ = cell[j];
            *t++ = cell[(j + 1) % 6];
            *t++ = (int)add;
        }
    }
}

// Spherical ECEF (reference :158-169), ENU (:178-201) and the FP32 scene
// vertex (:406-410) of every mesh vertex; the FP64 ENU point of each cell is
// kept in P for the frames.  Mesh vertex i < ncell is cell i, the others
// were written to P by k_mesh.
__global__ void k_ecef_enu(const double* __restrict__ clon, const double* __restrict__ clat,
                           const double* __restrict__ hsurf, int ncell, long long nmv,
                           EnuFrame E, double* __restrict__ P, float* __restrict__ vert) {
    const long long i = (long long)blockIdx.x * kBlockP + threadIdx.x;
    if (i >= nmv) return;
    double lon, lat, h;
    if (i < ncell) {
        lon = clon[i];
        lat = clat[i];
        h = hsurf[i];
    } else {
        lon = P[3 * i];
        lat = P[3 * i + 1];
        h = P[3 * i + 2];
    }
    const double R = kRadEarthP;
    double sl, cl, sp, cp;
    crt::sincos_cr(lon, &sl, &cl);
    crt::sincos_cr(lat, &sp, &cp);
    const double x = (R + h) * cp * cl;
    const double y = (R + h) * cp * sl;
    const double z = (R + h) * sp;
    const double so = E.so, co = E.co, sa = E.sa, ca = E.ca;
    const double dx = x - E.x0, dy = y - E.y0, dz = z - E.z0;
    const double r0 = -so * dx + co * dy;
    const double r1 = -sa * co * dx - sa * so * dy + ca * dz;
    const double r2 = ca * co * dx + ca * so * dy + sa * dz;
    if (i < ncell) {
        P[3 * i] = r0;
        P[3 * i + 1] = r1;
        P[3 * i + 2] = r2;
    }
    vert[3 * i] = (float)r0;
    vert[3 * i + 1] = (float)r1;
    vert[3 * i + 2] = (float)r2;
}

// Per-cell frame (reference :693-722): unit normal from the earth centre,
// unit north (pole direction minus its normal part), FP32 ray origin raised
// by ray_org_elev along the normal; flags cleared.
__global__ void k_frames(const double* __restrict__ P, int ncell, EnuFrame E,
                         double ray_org_elev, CellFrame* __restrict__ cells) {
    const int i = blockIdx.x * kBlockP + threadIdx.x;
    if (i >= ncell) return;
    const double p[3] = {P[3 * (size_t)i], P[3 * (size_t)i + 1], P[3 * (size_t)i + 2]};
    double n[3] = {p[0] - E.centre[0], p[1] - E.centre[1], p[2] - E.centre[2]};
    const double m = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= m;
    n[1] /= m;
    n[2] /= m;
    const double vn[3] = {E.pole[0] - p[0], E.pole[1] - p[1], E.pole[2] - p[2]};
    const double dp = vn[0] * n[0] + vn[1] * n[1] + vn[2] * n[2];
    double N[3] = {vn[0] - dp * n[0], vn[1] - dp * n[1], vn[2] - dp * n[2]};
    const double mN = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    N[0] /= mN;
    N[1] /= mN;
    N[2] /= mN;
    CellFrame f;
    for (int a = 0; a < 3; a++) {
        f.n[a] = n[a];
        f.north[a] = N[a];
        f.org[a] = (float)(p[a] + n[a] * ray_org_elev);
    }
    f.pad = 0;
    cells[i] = f;
}

// Flag bit 0 of every cell that is a corner of an ICON vertex with fewer
// than six cells (next to a hole of the mesh), as the host prologue does.
__global__ void k_hole_flags(const int* __restrict__ cov, int nvert, int ncell,
                             CellFrame* __restrict__ cells) {
    const int v = blockIdx.x * kBlockP + threadIdx.x;
    if (v >= nvert) return;
    int n = 0;
    for (int j = 0; j < 6; j++)
        if (cov[(size_t)nvert * j + v] != -2) n++;
    if (n == 6) return;
    for (int j = 0; j < 6; j++) {
        const int c = cov[(size_t)nvert * j + v];
        if (c >= 0 && c < ncell) atomicOr(&cells[c].pad, 1);
    }
}

// order-preserving map of a float to an unsigned key (min / max by integer
// atomics) and back
__device__ __forceinline__ unsigned fkey(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float fkey_inv(unsigned k) {
    const unsigned u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__global__ void k_bounds_init(unsigned* b) {
    if (threadIdx.x == 0) {
        b[0] = 0xffffffffu;
        b[1] = 0xffffffffu;
        b[2] = 0u;
        b[3] = 0u;
    }
}

// x / y extent of the FP32 vertices (min x, min y, max x, max y): a
// grid-stride pass, wavefront and workgroup reductions, one atomic per
// workgroup and value
__global__ __launch_bounds__(kBlockP) void k_bounds_xy(const float* __restrict__ vert,
                                                       long long nmv, unsigned* b) {
    unsigned kx0 = 0xffffffffu, ky0 = 0xffffffffu, kx1 = 0u, ky1 = 0u;
    for (long long i = (long long)blockIdx.x * kBlockP + threadIdx.x; i < nmv;
         i += (long long)gridDim.x * kBlockP) {
        const unsigned kx = fkey(vert[3 * i]), ky = fkey(vert[3 * i + 1]);
        kx0 = min(kx0, kx);
        ky0 = min(ky0, ky);
        kx1 = max(kx1, kx);
        ky1 = max(ky1, ky);
    }
    for (int o = 32; o > 0; o >>= 1) {
        kx0 = min(kx0, (unsigned)__shfl_xor((int)kx0, o, 64));
        ky0 = min(ky0, (unsigned)__shfl_xor((int)ky0, o, 64));
        kx1 = max(kx1, (unsigned)__shfl_xor((int)kx1, o, 64));
        ky1 = max(ky1, (unsigned)__shfl_xor((int)ky1, o, 64));
    }
    __shared__ unsigned part[kBlockP / 64][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[w][0] = kx0;
        part[w][1] = ky0;
        part[w][2] = kx1;
        part[w][3] = ky1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlockP / 64; k++) {
            kx0 = min(kx0, part[k][0]);
            ky0 = min(ky0, part[k][1]);
            kx1 = max(kx1, part[k][2]);
            ky1 = max(ky1, part[k][3]);
        }
        atomicMin(&b[0], kx0);
        atomicMin(&b[1], ky0);
        atomicMax(&b[2], kx1);
        atomicMax(&b[3], ky1);
    }
}

__global__ void k_sincos(const double* __restrict__ x, long long n, double* __restrict__ s,
                         double* __restrict__ c) {
    const long long i = (long long)blockIdx.x * kBlockP + threadIdx.x;
    if (i < n) crt::sincos_cr(x[i], &s[i], &c[i]);
}

bool args_in_range(const double* a, long long n) {
    const double lim = 1048576.0;  // 2^20: the exact range of the reduction
    for (long long i = 0; i < n; i++)
        if (!(std::fabs(a[i]) < lim)) return false;
    return true;
}

}  // namespace

#define PCHK(call)                                                              \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            err = std::string(#call) + ": " + hipGetErrorString(e_);           \
            rc = e_ == hipErrorOutOfMemory ? -3 : -2;                           \
            goto done;                                                          \
        }                                                                       \
    } while (0)

int prologue_device(const double* clon, const double* clat, const double* hsurf,
                    const double* vlon, const double* vlat, const int* cov, int ncell,
                    int nvert, int grid_type, double ray_org_elev, hipStream_t st,
                    Stager& sg, DeviceScene& out, std::string& err) {
    if (ncell < 1 || nvert < 0) {
        err = "num_cell must be >= 1 and num_vertex >= 0";
        return -1;
    }
    if ((long long)nvert * 6 > (long long)kLeafFirstMask) {
        err = "mesh has more triangles than the library supports (2^28)";
        return -1;
    }
    int rc = 0;
    // host side, concurrent with the device work below: the argument range of
    // the device sin / cos and the ENU origin (glibc sin / cos, cell order)
    EnuFrame E{};
    bool in_range = true;
    std::thread host([&]() {
        in_range = args_in_range(clon, ncell) && args_in_range(clat, ncell) &&
                   args_in_range(vlon, nvert) && args_in_range(vlat, nvert);
        enu_frame_host(clon, clat, hsurf, ncell, E);
    });
    double *d_in = nullptr, *d_P = nullptr;
    int *d_cov = nullptr, *d_cnt = nullptr;
    unsigned* d_b = nullptr;
    void* d_tmp = nullptr;
    size_t tmp_bytes = 0;
    int tot[2] = {0, 0};
    unsigned bk[4] = {0, 0, 0, 0};
    const size_t nv1 = (size_t)nvert + 1;
    double *d_clon, *d_clat, *d_hsurf, *d_vlon, *d_vlat;
    int *ntri_v, *nadd_v, *toff, *aoff;
    long long nmv = 0;
    PCHK(hipMalloc(&d_in, ((size_t)3 * ncell + (size_t)2 * nvert) * sizeof(double)));
    d_clon = d_in;
    d_clat = d_clon + ncell;
    d_hsurf = d_clat + ncell;
    d_vlon = d_hsurf + ncell;
    d_vlat = d_vlon + nvert;
    PCHK(hipMalloc(&d_cov, std::max<size_t>((size_t)6 * nvert, 1) * sizeof(int)));
    PCHK(hipMalloc(&d_cnt, 4 * nv1 * sizeof(int)));
    ntri_v = d_cnt;
    nadd_v = ntri_v + nv1;
    toff = nadd_v + nv1;
    aoff = toff + nv1;
    PCHK(hipMalloc(&d_b, 4 * sizeof(unsigned)));
    PCHK(sg.copy(d_clon, clon, (size_t)ncell * sizeof(double), true, st));
    PCHK(sg.copy(d_clat, clat, (size_t)ncell * sizeof(double), true, st));
    PCHK(sg.copy(d_hsurf, hsurf, (size_t)ncell * sizeof(double), true, st));
    if (nvert > 0) {
        PCHK(sg.copy(d_vlon, vlon, (size_t)nvert * sizeof(double), true, st));
        PCHK(sg.copy(d_vlat, vlat, (size_t)nvert * sizeof(double), true, st));
        PCHK(sg.copy(d_cov, cov, (size_t)6 * nvert * sizeof(int), true, st));
    }
    // triangle and added-vertex counts per ICON vertex, exclusive scans
    hipLaunchKernelGGL(k_valence, dim3(gridp((long long)nv1)), dim3(kBlockP), 0, st, d_cov,
                       nvert, grid_type, ntri_v, nadd_v);
    PCHK(rocprim::exclusive_scan(nullptr, tmp_bytes, ntri_v, toff, 0, nv1, rocprim::plus<int>(),
                                 st));
    PCHK(hipMalloc(&d_tmp, std::max<size_t>(tmp_bytes, 16)));
    PCHK(rocprim::exclusive_scan(d_tmp, tmp_bytes, ntri_v, toff, 0, nv1, rocprim::plus<int>(),
                                 st));
    PCHK(rocprim::exclusive_scan(d_tmp, tmp_bytes, nadd_v, aoff, 0, nv1, rocprim::plus<int>(),
                                 st));
    PCHK(hipMemcpyAsync(&tot[0], toff + nvert, sizeof(int), hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(&tot[1], aoff + nvert, sizeof(int), hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    out.ntri = tot[0];
    nmv = (long long)ncell + tot[1];
    out.nmv = nmv;
    PCHK(hipMalloc(&out.tri, std::max<size_t>((size_t)out.ntri * 3, 3) * sizeof(int)));
    PCHK(hipMalloc(&out.vert, std::max<size_t>((size_t)nmv * 3, 3) * sizeof(float)));
    PCHK(hipMalloc(&out.cells, (size_t)ncell * sizeof(CellFrame)));
    PCHK(hipMalloc(&d_P, (size_t)nmv * 3 * sizeof(double)));
    if (nvert > 0)
        hipLaunchKernelGGL(k_mesh, dim3(gridp(nvert)), dim3(kBlockP), 0, st, d_clon, d_clat,
                           d_hsurf, d_vlon, d_vlat, d_cov, nvert, ncell, grid_type, toff, aoff,
                           out.tri, d_P);
    PCHK(hipGetLastError());
    // the ENU constants are needed from here on
    host.join();
    if (!in_range) {
        err = "device prologue: longitudes and latitudes must be below 2^20 in magnitude";
        rc = -1;
        goto free_tmp;
    }
    out.lon0 = E.lon0;
    out.lat0 = E.lat0;
    hipLaunchKernelGGL(k_ecef_enu, dim3(gridp(nmv)), dim3(kBlockP), 0, st, d_clon, d_clat,
                       d_hsurf, ncell, nmv, E, d_P, out.vert);
    hipLaunchKernelGGL(k_frames, dim3(gridp(ncell)), dim3(kBlockP), 0, st, d_P, ncell, E,
                       ray_org_elev, out.cells);
    if (nvert > 0)
        hipLaunchKernelGGL(k_hole_flags, dim3(gridp(nvert)), dim3(kBlockP), 0, st, d_cov, nvert,
                           ncell, out.cells);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, d_b);
    hipLaunchKernelGGL(k_bounds_xy, dim3(std::min(gridp(nmv), 1024)), dim3(kBlockP), 0, st,
                       out.vert, nmv, d_b);
    PCHK(hipGetLastError());
    PCHK(hipMemcpyAsync(bk, d_b, sizeof(bk), hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    for (int k = 0; k < 4; k++) out.bounds_xy[k] = fkey_inv(bk[k]);
    goto free_tmp;
done:
    if (host.joinable()) host.join();
free_tmp:
    if (d_in) (void)hipFree(d_in);
    if (d_P) (void)hipFree(d_P);
    if (d_cov) (void)hipFree(d_cov);
    if (d_cnt) (void)hipFree(d_cnt);
    if (d_b) (void)hipFree(d_b);
    if (d_tmp) (void)hipFree(d_tmp);
    if (rc) {
        if (out.tri) (void)hipFree(out.tri);
        if (out.vert) (void)hipFree(out.vert);
        if (out.cells) (void)hipFree(out.cells);
        out.tri = nullptr;
        out.vert = nullptr;
        out.cells = nullptr;
    }
    return rc;
}

namespace {
// Flag bit 0 on the listed cells (the host's list of cells next to mesh holes).
__global__ void k_flag_cells(const int* __restrict__ list, int n, CellFrame* __restrict__ cells) {
    const int i = blockIdx.x * kBlockP + threadIdx.x;
    if (i < n) cells[list[i]].pad |= 1;
}
}  // namespace

void launch_flag_cells(const int* list, int n, CellFrame* cells, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(k_flag_cells, dim3(gridp(n)), dim3(kBlockP), 0, st, list, n, cells);
}

void launch_frames_from_enu(const double* P, int ncell, const double centre[3],
                            const double pole[3], double ray_org_elev, CellFrame* cells,
                            hipStream_t st) {
    EnuFrame E;
    std::memset(&E, 0, sizeof(E));
    for (int a = 0; a < 3; a++) {
        E.centre[a] = centre[a];
        E.pole[a] = pole[a];
    }
    hipLaunchKernelGGL(k_frames, dim3(gridp(ncell)), dim3(kBlockP), 0, st, P, ncell, E,
                       ray_org_elev, cells);
}

int sincos_cr_eval(const double* x, long long n, double* s, double* c, int device,
                   std::string& err) {
    if (n <= 0) return 0;
    if (!device) {
        for (long long i = 0; i < n; i++) crt::sincos_cr(x[i], &s[i], &c[i]);
        return 0;
    }
    int rc = 0;
    double* d = nullptr;
    PCHK(hipMalloc(&d, (size_t)n * 3 * sizeof(double)));
    PCHK(hipMemcpy(d, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_sincos, dim3(gridp(n)), dim3(kBlockP), 0, nullptr, d, n, d + n,
                       d + 2 * n);
    PCHK(hipGetLastError());
    PCHK(hipMemcpy(s, d + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    PCHK(hipMemcpy(c, d + 2 * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
done:
    if (d) (void)hipFree(d);
    return rc;
}

#undef PCHK

}  // namespace hz
