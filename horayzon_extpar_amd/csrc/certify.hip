// Walk-equivalence certificate of the lattice searches.
//
// The reference walk (mo_lradtopo_horayzon.cpp:521-546) carries its
// elevation from one azimuth to the next and steps 0.5 deg up or down until
// the occlusion answer flips; the lattice searches (climb, learned start)
// find the top hit index of each azimuth on its own.  Both give the same
// value wherever hits are monotone in elevation over the indices the walk
// explores.  Within one azimuth's vertical half-plane the mesh cuts the
// plane in polylines, and a polyline can only begin or end where the plane
// crosses a mesh edge with a single triangle (the domain border and the
// holes grid_type 1 leaves at vertices with fewer than six cells, :299).  A
// ray that misses while a higher ray hits must pass below the start of the
// polyline the higher ray hits, so rays above the elevation of every such
// crossing point see monotone hits.  The crossing point lies on an edge
// between two boundary points, so its tangent of elevation w / u (w along
// the cell normal, u along the azimuth direction) is at most the larger of
// the endpoints'; endpoints lie within one edge length lw of the plane.
//
// k_certify computes, per cell and azimuth, the tangent bound B = max w / u
// over boundary points in the band |lateral| <= lw, 0 < u <= reach + lw
// (B = +inf when such a point is within lw of the cell: the bound does not
// hold for crossings behind the origin), turns it into the lattice index
// bound kb, runs the backward pass that forces casts where a later azimuth
// needs to know where the walk ended, and simulates the walk's carried index
// forward over the lattice results.  Cells whose walk never has to be cast
// get their sector means and SVF from the lattice results here; the others
// are listed for the walk kernel (trace.hip, MODE 3).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hz_internal.h"

namespace hz {
namespace {

constexpr double kPi = 3.14159265358979323846;

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
    const double2* az;
    const double* tanv;
    double x0, y0, S;
    int nx, ny, rb, L;
    double lw, reach, r2max, shift;
};

// Tangent bound rounded up to float (a conservative, exactly reproducible
// storage of the running maximum).
__device__ __forceinline__ float up_float(double b) { return __double2float_ru(b); }

// Lattice index bound of a float tangent bound: the lowest lattice index
// whose tangent exceeds it by the margin (kMaxLattice: none below 90 deg).
__device__ __forceinline__ int kb_of(float bf, const double* __restrict__ tanv) {
    if (bf == -INFINITY) return kCertNone;
    if (!(bf < INFINITY)) return kMaxLattice;
    const double bm = (double)bf + kCertMargin;
    // first k in [kWalkFloor, kMaxLattice) with tan(k * 0.5 deg) > bm
    int lo = 0, hi = kCertTan;  // search over table positions
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (tanv[mid] > bm) hi = mid; else lo = mid + 1;
    }
    return lo + kWalkFloor;  // kMaxLattice when no tangent exceeds bm
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_certify(CertArgs C, const CellFrame* __restrict__ cells,
                                                   const int* __restrict__ order,
                                                   const double* __restrict__ hitem, WalkParams P,
                                                   double* __restrict__ horizon,
                                                   double* __restrict__ skyview,
                                                   int* __restrict__ kb, int* __restrict__ fixlist,
                                                   TraceState* __restrict__ ts, int skip_border) {
    // positions [range[0], range[0] + range[1]) of the trace order, or all
    const int pr = blockIdx.x * BLOCK + threadIdx.x;
    const int ncell = P.ncell;
    const int npos = P.range ? P.range[1] : ncell;
    const int p = (P.range ? P.range[0] : 0) + pr;
    bool fix = false;
    int cell = 0;
    if (pr < npos) cell = order ? order[p] : p;
    if (pr < npos && !(skip_border && (cells[cell].pad & 1))) {
        const int L = C.L;
        const CellFrame& f = cells[cell];
        const double ox = f.org[0], oy = f.org[1], oz = f.org[2];
        int bx = (int)floor((ox - C.x0) / C.S), by = (int)floor((oy - C.y0) / C.S);
        bx = min(max(bx, 0), C.nx - 1);
        by = min(max(by, 0), C.ny - 1);
        const bool near = C.near[by * C.nx + bx] != 0;
        if (near) {
            const double n0 = f.n[0], n1 = f.n[1], n2 = f.n[2];
            const double N0 = f.north[0], N1 = f.north[1], N2 = f.north[2];
            // east = north x normal: azimuth phi has the horizontal direction
            // cos(phi) north + sin(phi) east (clockwise from north)
            const double E0 = N1 * n2 - N2 * n1, E1 = N2 * n0 - N0 * n2, E2 = N0 * n1 - N1 * n0;
            float* B = reinterpret_cast<float*>(kb);
            for (int j = 0; j < L; j++) B[(size_t)j * ncell + p] = -INFINITY;
            const double ulim = C.reach + C.lw;
            const double dphi = (2.0 * kPi) / (double)L;
            for (int yy = max(0, by - C.rb); yy <= min(C.ny - 1, by + C.rb); yy++) {
                for (int xx = max(0, bx - C.rb); xx <= min(C.nx - 1, bx + C.rb); xx++) {
                    const int b = yy * C.nx + xx;
                    for (int q = C.off[b]; q < C.off[b + 1]; q++) {
                        // every boundary point counts, the cell's own vertex too:
                        // the FP32 origin lies up to about a centimetre off it
                        // horizontally, so a near-vertical ray can pass the mesh
                        // edge there (u <= lw below: no bound)
                        const float4 v = C.pts[q];
                        const double rx = (double)v.x - ox, ry = (double)v.y - oy,
                                     rz = (double)v.z - oz;
                        const double uN = rx * N0 + ry * N1 + rz * N2;
                        const double uE = rx * E0 + ry * E1 + rz * E2;
                        const double r2 = uN * uN + uE * uE;
                        if (r2 > C.r2max) continue;
                        const double w = rx * n0 + ry * n1 + rz * n2;
                        const double r = sqrt(r2);
                        // candidate azimuths: every one when the point is
                        // within about lw, else a window around its bearing
                        // wider than the band (the exact test follows)
                        int j0 = 0, nj = L;
                        if (r > 1.05 * C.lw) {
                            const float phi = atan2f((float)uE, (float)uN);
                            const float half = asinf((float)(C.lw / r)) + 0.02f;
                            const double rel = ((double)phi - C.shift) / dphi;
                            const double wj = (double)half / dphi;
                            j0 = (int)floor(rel - wj) - 1;
                            nj = min(L, (int)ceil(rel + wj) + 1 - j0 + 1);
                        }
                        for (int t = 0; t < nj; t++) {
                            int j = (j0 + t) % L;
                            if (j < 0) j += L;
                            const double2 cs = C.az[j];
                            const double u = uN * cs.x + uE * cs.y;
                            const double lat = uE * cs.x - uN * cs.y;
                            if (!(u > 0.0) || u > ulim || fabs(lat) > C.lw) continue;
                            const float tb = u <= C.lw ? INFINITY : up_float(w / u);
                            float* bj = B + (size_t)j * ncell + p;
                            if (tb > *bj) *bj = tb;
                        }
                    }
                }
            }
            // bounds -> lattice indices, then the backward pass: an azimuth
            // with no hit at index 0 ends at an unknown negative index unless
            // cast, so it is cast when the next azimuth needs a known start
            for (int j = 0; j < L; j++) {
                const size_t at = (size_t)j * ncell + p;
                kb[at] = kb_of(B[at], C.tanv);
            }
            bool nk_next = false;
            for (int j = L - 1; j >= 0; j--) {
                const size_t at = (size_t)j * ncell + p;
                int k = kb[at];
                const int top = lattice_top(hitem[at], P.hori_acc);
                const bool none = k == kCertNone;
                bool nk;
                if (top >= 0) {
                    nk = !none;
                } else {
                    if (nk_next) k = kCertAll;
                    nk = !none || nk_next;
                }
                kb[at] = k;
                nk_next = nk;
            }
            // forward: the walk's carried index over the lattice results
            bool known = true;
            int k = 0;
            for (int j = 0; j < L && !fix; j++) {
                const size_t at = (size_t)j * ncell + p;
                const int top = lattice_top(hitem[at], P.hori_acc);
                if (cert_cast(known, k, top, kb[at])) {
                    fix = true;
                } else {
                    cert_skip_state(known, k, top);
                }
            }
        }
        if (!fix) {
            // sector means and SVF from the lattice results, in the order of
            // the cell-by-cell search (:736-754)
            double sector = 0.0, svf = 0.0;
            int ksec = 0, jsec = 0;
            for (int j = 0; j < L; j++) {
                const double h = hitem[(size_t)j * ncell + p];
                sector += h;
                if (++ksec == P.refine_factor) {
                    horizon[(size_t)jsec * ncell + cell] =
                        ((sector / kPi) * 180.0) / (double)P.refine_factor;
                    sector = 0.0;
                    ksec = 0;
                    jsec++;
                }
                double sn = 0.0;
                if (h > 0.0) {
                    int kk = (int)rint(h / (2.0 * P.hori_acc) - 0.5);
                    kk = kk < 0 ? 0 : (kk > kMaxLattice ? kMaxLattice : kk);
                    sn = P.sin_h[kk];
                }
                if (P.svf_type == 0)
                    svf += (1.0 - sn);
                else if (P.svf_type == 1)
                    svf += (1.0 - (sn * sn));
                else
                    svf += (1.0 - (sn * sn * sn));
            }
            skyview[cell] = svf / (double)L;
        }
    }
    // list the cells for the walk: one atomic per wavefront, lanes in order
    const unsigned long long m = __ballot(fix);
    if (m) {
        const int lane = threadIdx.x & 63;
        int base = 0;
        if (lane == __ffsll((long long)m) - 1) base = atomicAdd(&ts->nfix, __popcll(m));
        base = __shfl(base, __ffsll((long long)m) - 1, 64);
        if (fix) fixlist[base + __popcll(m & ((1ull << lane) - 1ull))] = p;
    }
}

__global__ void k_zero_ints(int* __restrict__ p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

void cert_tables(int num_hori, int refine_factor, double2* az, double* tanv,
                 double* shift_out) {
    const int L = num_hori * refine_factor;
    const double d2r = kPi / 180.0;
    double shift = 0.0;
    if (refine_factor != 1)
        shift = -((360.0 * d2r) / (2.0 * num_hori)) + ((360.0 * d2r) / (2.0 * (double)L));
    const double dphi = (360.0 * d2r) / (double)L;
    if (shift_out) *shift_out = shift;
    for (int j = 0; j < L; j++) {
        double s, c;
        ::sincos(shift + (double)j * dphi, &s, &c);
        az[j] = make_double2(c, s);
    }
    for (int k = kWalkFloor; k < kMaxLattice; k++) tanv[k - kWalkFloor] = ::tan((double)k * (0.5 * d2r));
}

void cert_index_free(CertIndex& ci) {
    void* b[] = {ci.pts, ci.off, ci.near, ci.az, ci.tanv};
    for (void* x : b)
        if (x) (void)hipFree(x);
    ci = CertIndex();
}

int cert_index_build(CertIndex& ci, const CellFrame* cells, const float* vert, int ncell,
                     const int* tri, long long ntri, const float bounds_xy[4], double hmin,
                     double hmax, float tfar, int num_hori, int refine_factor, hipStream_t st,
                     std::string& err) {
    const int L = num_hori * refine_factor;
    if (ci.az && ci.dist_key == tfar && ci.az_key_nhori == num_hori &&
        ci.az_key_rf == refine_factor)
        return 0;
    // azimuth / tangent tables (depend on nhori, rf)
    if (ci.az_key_nhori != num_hori || ci.az_key_rf != refine_factor || !ci.az) {
        if (ci.az) HZ_HIP(hipFree(ci.az), err);
        if (ci.tanv) HZ_HIP(hipFree(ci.tanv), err);
        ci.az = nullptr;
        ci.tanv = nullptr;
        std::vector<double2> az(L);
        std::vector<double> tv(kCertTan);
        cert_tables(num_hori, refine_factor, az.data(), tv.data(), &ci.shift);
        HZ_HIP(hipMalloc(&ci.az, (size_t)L * sizeof(double2)), err);
        HZ_HIP(hipMalloc(&ci.tanv, (size_t)kCertTan * sizeof(double)), err);
        HZ_HIP(hipMemcpyAsync(ci.az, az.data(), az.size() * sizeof(double2),
                              hipMemcpyHostToDevice, st), err);
        HZ_HIP(hipMemcpyAsync(ci.tanv, tv.data(), tv.size() * sizeof(double),
                              hipMemcpyHostToDevice, st), err);
        HZ_HIP(hipStreamSynchronize(st), err);  // the host tables go out of scope
        ci.L = L;
        ci.az_key_nhori = num_hori;
        ci.az_key_rf = refine_factor;
    }
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
        // lateral band: the longest boundary edge, with a relative and an
        // absolute margin
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

// Flag bit 1 of every cell (bit 0 clear) that has a boundary point in reach
// of some azimuth's band (the test of k_certify without the azimuths), or
// clear it (clear_only); counts[0] / counts[1] receive the numbers of cells
// with bit 0 and with bit 1 only.
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

// Positions of the cells next to a mesh hole, wave-aggregated appends.
__global__ void k_border_list(const CellFrame* __restrict__ cells, const int* __restrict__ order,
                              int ncell, int* __restrict__ list, TraceState* __restrict__ ts) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool f = p < ncell && (cells[order ? order[p] : p].pad & 1);
    const unsigned long long m = __ballot(f);
    if (m) {
        const int lane = threadIdx.x & 63;
        const int first = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane == first) base = atomicAdd(&ts->nfix, __popcll(m));
        base = __shfl(base, first, 64);
        if (f) list[base + __popcll(m & ((1ull << lane) - 1ull))] = p;
    }
}

void launch_border_list(const CellFrame* cells, const int* order, int ncell, int* list,
                        TraceState* ts, hipStream_t st) {
    hipLaunchKernelGGL(k_border_list, dim3((ncell + 255) / 256), dim3(256), 0, st, cells, order,
                       ncell, list, ts);
}

static CertArgs cert_args(const CertIndex& ci) {
    CertArgs C;
    C.pts = ci.pts;
    C.off = ci.off;
    C.near = ci.near;
    C.az = ci.az;
    C.tanv = ci.tanv;
    C.x0 = ci.x0;
    C.y0 = ci.y0;
    C.S = ci.S;
    C.nx = ci.nx;
    C.ny = ci.ny;
    C.rb = ci.rb;
    C.L = ci.L;
    C.lw = ci.lw;
    C.reach = ci.reach;
    C.r2max = ci.r2max;
    C.shift = ci.shift;
    return C;
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

void launch_certify(const CertIndex& ci, const CellFrame* cells, const int* order,
                    const double* hitem, const WalkParams& p, double* horizon, double* skyview,
                    int* kb, int* fixlist, TraceState* ts, bool skip_border, hipStream_t st) {
    const CertArgs C = cert_args(ci);
    hipLaunchKernelGGL(k_certify<128>, dim3((p.ncell + 127) / 128), dim3(128), 0, st, C, cells,
                       order, hitem, p, horizon, skyview, kb, fixlist, ts, skip_border ? 1 : 0);
}

}  // namespace hz
