// Where the lattice searches equal the reference walk.
//
// The reference walk (mo_lradtopo_horayzon.cpp:521-546) carries its
// elevation from one azimuth to the next and steps 0.5 deg up or down until
// the occlusion answer flips; the lattice searches (climb, learned start)
// find the top hit index of each azimuth on its own.  Both give the same
// value wherever hits are monotone in elevation.  Within one azimuth's
// vertical half-plane the mesh cuts the plane in polylines, and a polyline
// can only begin where the plane crosses a mesh edge with a single triangle
// (the domain border and the holes grid_type 1 leaves at vertices with
// fewer than six cells, :299).  A ray that misses while a higher ray hits
// must pass below the start of the polyline the higher ray hits, so a cell
// with no such crossing within reach of its rays sees monotone hits in every
// azimuth.  Crossings lie on edges between boundary points (cells next to
// such a vertex), within one edge length of either end, so a cell with no
// boundary point within reach + the longest boundary edge (horizontal
// distance in its own frame) is such a cell: a "far" cell.
//
// cert_index_build gathers the boundary points into a bucket grid;
// k_cert_near flags the other cells ("near", CellFrame flag bit 1) and
// counts the classes, which the trace order sorts by: cells next to mesh
// holes, near cells, far cells.  The certified searches run the reference
// walk on the first two classes and the lattice search on the far cells.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hz_internal.h"

namespace hz {
namespace {

// Boundary points (cells with flag bit 0) and the longest triangle edge
// between two of them.
__global__ void k_cert_gather(const CellFrame* __restrict__ cells, const float* __restrict__ vert,
                              int ncell, const int* __restrict__ tri, long long ntri,
                              float4* __restrict__ pts, unsigned* __restrict__ cnt) {
    const long long n = ncell > ntri ? (long long)ncell : ntri;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < ncell && (cells[i].pad & 1)) {
            const unsigned slot = atomicAdd(&cnt[0], 1u);
            pts[slot] = make_float4(vert[3 * i], vert[3 * i + 1], vert[3 * i + 2], 0.0f);
        }
        if (i < ntri) {
            // longest triangle edge between two boundary points (boundary
            // edges have both ends there)
            float best = 0.0f;
            for (int e = 0; e < 3; e++) {
                const int a = tri[3 * i + e], b = tri[3 * i + (e + 1) % 3];
                if (a >= ncell || b >= ncell) continue;
                if (!(cells[a].pad & 1) || !(cells[b].pad & 1)) continue;
                const float dx = vert[3 * b] - vert[3 * a];
                const float dy = vert[3 * b + 1] - vert[3 * a + 1];
                const float dz = vert[3 * b + 2] - vert[3 * a + 2];
                best = fmaxf(best, sqrtf(dx * dx + dy * dy + dz * dz));
            }
            if (best > 0.0f) atomicMax(&cnt[1], __float_as_uint(best));
        }
    }
}

struct CertArgs {
    const float4* pts;
    const int* off;
    const unsigned char* near;
    double x0, y0, S;
    int nx, ny, rb;
    double r2max;
};

CertArgs cert_args(const CertIndex& ci) {
    CertArgs C;
    C.pts = ci.pts;
    C.off = ci.off;
    C.near = ci.near;
    C.x0 = ci.x0;
    C.y0 = ci.y0;
    C.S = ci.S;
    C.nx = ci.nx;
    C.ny = ci.ny;
    C.rb = ci.rb;
    C.r2max = ci.r2max;
    return C;
}

__global__ void k_zero_ints(int* __restrict__ p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Flag bit 1 of every cell (bit 0 clear) with a boundary point within
// horizontal distance sqrt(r2max) in its own frame, or clear it
// (clear_only); counts[0] / counts[1] receive the numbers of cells with bit
// 0 and with bit 1 only.
__global__ void k_cert_near(CertArgs C, CellFrame* __restrict__ cells, int ncell, int clear_only,
                            int* __restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int border = 0, near = 0;
    if (i < ncell) {
        CellFrame& f = cells[i];
        int fl = f.pad & ~2;
        border = fl & 1;
        if (!clear_only && !border) {
            const double ox = f.org[0], oy = f.org[1], oz = f.org[2];
            int bx = (int)floor((ox - C.x0) / C.S), by = (int)floor((oy - C.y0) / C.S);
            bx = min(max(bx, 0), C.nx - 1);
            by = min(max(by, 0), C.ny - 1);
            if (C.near[by * C.nx + bx]) {
                const double n0 = f.n[0], n1 = f.n[1], n2 = f.n[2];
                const double N0 = f.north[0], N1 = f.north[1], N2 = f.north[2];
                const double E0 = N1 * n2 - N2 * n1, E1 = N2 * n0 - N0 * n2,
                             E2 = N0 * n1 - N1 * n0;
                for (int yy = max(0, by - C.rb); yy <= min(C.ny - 1, by + C.rb) && !near; yy++)
                    for (int xx = max(0, bx - C.rb); xx <= min(C.nx - 1, bx + C.rb) && !near; xx++) {
                        const int b = yy * C.nx + xx;
                        for (int q = C.off[b]; q < C.off[b + 1]; q++) {
                            const float4 v = C.pts[q];
                            const double rx = (double)v.x - ox, ry = (double)v.y - oy,
                                         rz = (double)v.z - oz;
                            const double uN = rx * N0 + ry * N1 + rz * N2;
                            const double uE = rx * E0 + ry * E1 + rz * E2;
                            if (uN * uN + uE * uE <= C.r2max) {
                                near = 1;
                                break;
                            }
                        }
                    }
            }
        }
        if (near) fl |= 2;
        f.pad = fl;
    }
    const unsigned long long mb = __ballot(border != 0), mn = __ballot(near != 0);
    if ((threadIdx.x & 63) == 0) {
        if (mb) atomicAdd(&counts[0], __popcll(mb));
        if (mn) atomicAdd(&counts[1], __popcll(mn));
    }
}

// Position ranges of the three classes in the trace order (border cells,
// cells with a boundary point in reach, the rest), then of the last two
// together: {first, count} each.
__global__ void k_cert_ranges(const int* __restrict__ counts, int ncell, int* __restrict__ ranges) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int nb = counts[0], nn = counts[1];
        ranges[0] = 0;
        ranges[1] = nb;
        ranges[2] = nb;
        ranges[3] = nn;
        ranges[4] = nb + nn;
        ranges[5] = ncell - nb - nn;
        ranges[6] = nb;  // near and far together (the lattice items)
        ranges[7] = ncell - nb;
    }
}

}  // namespace

void cert_index_free(CertIndex& ci) {
    void* b[] = {ci.pts, ci.off, ci.near};
    for (void* x : b)
        if (x) (void)hipFree(x);
    ci = CertIndex();
}

int cert_index_build(CertIndex& ci, const CellFrame* cells, const float* vert, int ncell,
                     const int* tri, long long ntri, const float bounds_xy[4], double hmin,
                     double hmax, float tfar, hipStream_t st, std::string& err) {
    if (ci.dist_key == tfar && ci.off) return 0;
    if (ci.dist_key != tfar || !ci.off) {
        // boundary points and the longest boundary edge, gathered on the device
        float4* tmp = nullptr;
        unsigned* cnt = nullptr;
        HZ_HIP(hipMalloc(&tmp, std::max(1, ncell) * sizeof(float4)), err);
        if (hipMalloc(&cnt, 2 * sizeof(unsigned)) != hipSuccess) {
            (void)hipFree(tmp);
            err = "hipMalloc failed (certificate index)";
            return -3;
        }
        std::vector<float4> pts;
        unsigned hc[2] = {0u, 0u};
        hipError_t e = hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned), st);
        if (e == hipSuccess) {
            const long long n = std::max<long long>(ncell, ntri);
            const int grid = (int)std::min<long long>(4096, (n + 255) / 256);
            hipLaunchKernelGGL(k_cert_gather, dim3(std::max(1, grid)), dim3(256), 0, st, cells, vert,
                               ncell, tri, ntri, tmp, cnt);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess && hc[0] > 0) {
            pts.resize(hc[0]);
            e = hipMemcpyAsync(pts.data(), tmp, pts.size() * sizeof(float4),
                               hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
        }
        (void)hipFree(tmp);
        (void)hipFree(cnt);
        if (e != hipSuccess) {
            err = std::string("certificate index: ") + hipGetErrorString(e);
            return -2;
        }
        float maxlen = 0.0f;
        std::memcpy(&maxlen, &hc[1], sizeof(float));
        // the longest boundary edge, with a relative and an absolute margin
        ci.lw = (double)maxlen * 1.001 + 1.0;
        ci.reach = (double)tfar;
        const double ulim = ci.reach + ci.lw;
        ci.r2max = ulim * ulim + ci.lw * ci.lw;
        // a boundary point in reach is within this x/y distance of the cell
        // origin: horizontal distance in the cell's frame, plus the vertical
        // offset (terrain range, origin height, earth curvature over the reach)
        const double dz = (hmax - hmin) + 10.0 + ulim * ulim / (2.0 * 6371229.0);
        const double rxy = std::sqrt(ci.r2max + dz * dz) * 1.0001 + 1.0;
        // bucket grid over the scene's x/y extent
        double w = std::max(1.0, (double)bounds_xy[2] - bounds_xy[0]);
        double hgt = std::max(1.0, (double)bounds_xy[3] - bounds_xy[1]);
        double S = std::max(rxy / 6.0, std::max(w, hgt) / 2048.0);
        ci.x0 = (double)bounds_xy[0] - S;
        ci.y0 = (double)bounds_xy[1] - S;
        ci.nx = (int)std::ceil((w + 2.0 * S) / S) + 1;
        ci.ny = (int)std::ceil((hgt + 2.0 * S) / S) + 1;
        ci.S = S;
        // a point in bucket b2 is within rxy of a cell in bucket b only if the
        // bucket centres are within rxy + S * sqrt(2)
        const double rcen = rxy + S * 1.4142135623730951;
        ci.rb = (int)std::ceil(rcen / S);
        const int nb = ci.nx * ci.ny;
        std::vector<int> off(nb + 1, 0), bucket(pts.size());
        for (size_t i = 0; i < pts.size(); i++) {
            int bx = (int)std::floor(((double)pts[i].x - ci.x0) / S);
            int by = (int)std::floor(((double)pts[i].y - ci.y0) / S);
            bx = std::min(std::max(bx, 0), ci.nx - 1);
            by = std::min(std::max(by, 0), ci.ny - 1);
            bucket[i] = by * ci.nx + bx;
            off[bucket[i] + 1]++;
        }
        for (int b = 0; b < nb; b++) off[b + 1] += off[b];
        std::vector<float4> sorted(pts.size());
        {
            std::vector<int> fill(off.begin(), off.end() - 1);
            for (size_t i = 0; i < pts.size(); i++) sorted[fill[bucket[i]]++] = pts[i];
        }
        // near flags: any non-empty bucket whose centre is within rcen
        std::vector<unsigned char> nearf(nb, 0);
        for (int by = 0; by < ci.ny; by++)
            for (int bx = 0; bx < ci.nx; bx++) {
                unsigned char any = 0;
                for (int yy = std::max(0, by - ci.rb); yy <= std::min(ci.ny - 1, by + ci.rb) && !any; yy++)
                    for (int xx = std::max(0, bx - ci.rb); xx <= std::min(ci.nx - 1, bx + ci.rb); xx++) {
                        const int b2 = yy * ci.nx + xx;
                        if (off[b2 + 1] == off[b2]) continue;
                        const double ddx = (double)(xx - bx) * S, ddy = (double)(yy - by) * S;
                        if (ddx * ddx + ddy * ddy <= rcen * rcen) {
                            any = 1;
                            break;
                        }
                    }
                nearf[by * ci.nx + bx] = any;
            }
        if (ci.pts) HZ_HIP(hipFree(ci.pts), err);
        if (ci.off) HZ_HIP(hipFree(ci.off), err);
        if (ci.near) HZ_HIP(hipFree(ci.near), err);
        ci.pts = nullptr;
        ci.off = nullptr;
        ci.near = nullptr;
        ci.npts = (int)pts.size();
        HZ_HIP(hipMalloc(&ci.pts, std::max<size_t>(1, sorted.size()) * sizeof(float4)), err);
        HZ_HIP(hipMalloc(&ci.off, off.size() * sizeof(int)), err);
        HZ_HIP(hipMalloc(&ci.near, nearf.size()), err);
        if (!sorted.empty())
            HZ_HIP(hipMemcpyAsync(ci.pts, sorted.data(), sorted.size() * sizeof(float4),
                                  hipMemcpyHostToDevice, st), err);
        HZ_HIP(hipMemcpyAsync(ci.off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice,
                              st), err);
        HZ_HIP(hipMemcpyAsync(ci.near, nearf.data(), nearf.size(), hipMemcpyHostToDevice, st), err);
        HZ_HIP(hipStreamSynchronize(st), err);
        ci.dist_key = tfar;
    }
    return 0;
}

void launch_cert_classes(const CertIndex& ci, CellFrame* cells, int ncell, bool clear_only,
                         int* counts_ranges, hipStream_t st) {
    // counts_ranges: 2 counters then 6 range entries
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, counts_ranges, 2);
    hipLaunchKernelGGL(k_cert_near, dim3((ncell + 255) / 256), dim3(256), 0, st, cert_args(ci),
                       cells, ncell, clear_only ? 1 : 0, counts_ranges);
    hipLaunchKernelGGL(k_cert_ranges, dim3(1), dim3(64), 0, st, counts_ranges, ncell,
                       counts_ranges + 2);
}

}  // namespace hz
