// Per-cell descriptor of the learned elevation bracket: for each of 24
// azimuths and 8 log-spaced radial rings (0.25 .. 40 km), the largest
// elevation angle of the terrain seen from the cell's ray origin, sampled on
// a max-height raster of the FP32 scene.  (No reference counterpart.)
//
//   k_raster_clear / k_raster_tris : every triangle writes the linear
//       interpolation of its vertex heights at the pixel centres it covers
//       (atomicMax of order-preserving encoded floats)
//   k_descriptor : one lane per cell, 24 x 8 x kSamples raster samples; the
//       elevation angle uses the cell's own up vector, so earth curvature and
//       the tilt of distant cells' frames are taken into account.
//
// The sample distances and the azimuth directions are the same for every
// cell; the host evaluates them in FP64 and rounds them to FP32 (DescTables),
// so every other step is plain FP32 arithmetic that a CPU restatement
// repeats exactly.

#include <cmath>

#include "hz_internal.h"
#include "mlp.h"

namespace hz {

namespace {

constexpr int kSamples = 3;  // samples per ring
constexpr int kSamplesAz = kDescRings * kSamples;  // samples per azimuth

// sample s (0-based) of ring k at dist[k * kSamples + s]; azimuth j
// (clockwise from north) has direction cos[j] north + sin[j] east
struct DescTables {
    float dist[kSamplesAz];
    float sin[kDescAz];
    float cos[kDescAz];
};

__device__ __forceinline__ unsigned f2ord(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ uint16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr unsigned kEmpty = 0x007fffffu;  // f2ord(-inf) region: below any height

__global__ void k_raster_clear(unsigned* __restrict__ z, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = kEmpty;
}

__global__ void k_raster_tris(const float* __restrict__ vert, const int* __restrict__ tri, int ntri,
                              unsigned* __restrict__ z, int nx, int ny, float x0, float y0,
                              float inv_pix) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntri) return;
    float px[3], py[3], pz[3];
    for (int k = 0; k < 3; k++) {
        int v = tri[3 * t + k];
        px[k] = (vert[3 * v] - x0) * inv_pix;
        py[k] = (vert[3 * v + 1] - y0) * inv_pix;
        pz[k] = vert[3 * v + 2];
    }
    float det = (px[1] - px[0]) * (py[2] - py[0]) - (px[2] - px[0]) * (py[1] - py[0]);
    if (det == 0.0f) return;
    float idet = 1.0f / det;
    int ix0 = max(0, (int)ceilf(fminf(px[0], fminf(px[1], px[2]))));
    int ix1 = min(nx - 1, (int)floorf(fmaxf(px[0], fmaxf(px[1], px[2]))));
    int iy0 = max(0, (int)ceilf(fminf(py[0], fminf(py[1], py[2]))));
    int iy1 = min(ny - 1, (int)floorf(fmaxf(py[0], fmaxf(py[1], py[2]))));
    // bound the loop (a sliver triangle spanning the raster would be absurd)
    if ((ix1 - ix0) > 256 || (iy1 - iy0) > 256) return;
    for (int iy = iy0; iy <= iy1; iy++)
        for (int ix = ix0; ix <= ix1; ix++) {
            float qx = (float)ix, qy = (float)iy;
            float b1 = ((qx - px[0]) * (py[2] - py[0]) - (px[2] - px[0]) * (qy - py[0])) * idet;
            float b2 = ((px[1] - px[0]) * (qy - py[0]) - (qx - px[0]) * (py[1] - py[0])) * idet;
            float b0 = 1.0f - b1 - b2;
            const float e = -1e-4f;
            if (b0 < e || b1 < e || b2 < e) continue;
            float h = b0 * pz[0] + b1 * pz[1] + b2 * pz[2];
            atomicMax(&z[(size_t)iy * nx + ix], f2ord(h));
        }
}

__global__ __launch_bounds__(128) void k_descriptor(const unsigned* __restrict__ z, int nx, int ny,
                                                    float x0, float y0, float inv_pix,
                                                    const CellFrame* __restrict__ cells, int ncell,
                                                    DescTables tab, uint16_t* __restrict__ desc) {
    const int c = blockIdx.x * 128 + threadIdx.x;
    if (c >= ncell) return;
    const CellFrame f = cells[c];
    const float ox = f.org[0], oy = f.org[1], oz = f.org[2];
    const float nx3 = (float)f.n[0], ny3 = (float)f.n[1], nz3 = (float)f.n[2];
    const float Nx = (float)f.north[0], Ny = (float)f.north[1], Nz = (float)f.north[2];
    // east = north x n
    const float Ex = Ny * nz3 - Nz * ny3, Ey = Nz * nx3 - Nx * nz3;
    for (int j = 0; j < kDescAz; j++) {
        const float sphi = tab.sin[j], cphi = tab.cos[j];
        const float hx = cphi * Nx + sphi * Ex, hy = cphi * Ny + sphi * Ey;
        // the azimuth's raster reads are issued together (samples off the
        // raster read pixel 0 and are dropped): they are the kernel's latency
        unsigned zu[kSamplesAz];
#pragma unroll
        for (int i = 0; i < kSamplesAz; i++) {
            const float d = tab.dist[i];
            const float qx = ox + d * hx, qy = oy + d * hy;
            const int ix = (int)floorf((qx - x0) * inv_pix + 0.5f);
            const int iy = (int)floorf((qy - y0) * inv_pix + 0.5f);
            const bool in = ix >= 0 && iy >= 0 && ix < nx && iy < ny;
            zu[i] = z[in ? (size_t)iy * nx + ix : 0];
            if (!in) zu[i] = kEmpty;
        }
        uint16_t* dst = desc + (size_t)c * kIn + j * kDescRings;
#pragma unroll
        for (int k = 0; k < kDescRings; k++) {
            // the sample of largest elevation atan2(vn, hor), hor > 0, is the
            // one of largest vn / hor: compared as vn^2 hor_b^2 against
            // vn_b^2 hor^2 by sign, so atan2f runs once per bin
            bool any = false;
            float bn = 0.0f, bq = 1.0f;  // vn and hor^2 of the best sample
#pragma unroll
            for (int s = 0; s < kSamples; s++) {
                const unsigned u = zu[k * kSamples + s];
                if (u == kEmpty) continue;
                const float d = tab.dist[k * kSamples + s];
                const float vx = (ox + d * hx) - ox, vy = (oy + d * hy) - oy, vz = ord2f(u) - oz;
                const float vn = vx * nx3 + vy * ny3 + vz * nz3;
                const float q = fmaxf(vx * vx + vy * vy + vz * vz - vn * vn, 1e-6f);
                bool better = !any;
                if (any) {
                    if ((vn >= 0.0f) != (bn >= 0.0f)) {
                        better = vn >= 0.0f;
                    } else {
                        const float l = vn * vn * bq, rr = bn * bn * q;
                        better = vn >= 0.0f ? l > rr : l < rr;
                    }
                }
                if (better) {
                    bn = vn;
                    bq = q;
                }
                any = true;
            }
            const float best = any ? atan2f(bn, sqrtf(bq)) * 57.29577951308232f : -90.0f;
            dst[k] = f2bf(best * 0.1f);
        }
    }
}

}  // namespace

int Raster::reserve(int n, std::string& err) {
    if (n <= cap) return 0;
    free_all();
    HZ_HIP(hipMalloc(&z, (size_t)n * sizeof(unsigned)), err);
    cap = n;
    return 0;
}

void Raster::free_all() {
    if (z) (void)hipFree(z);
    z = nullptr;
    cap = 0;
}

// bounds_host: xmin, ymin, xmax, ymax of the scene [m]
int raster_build(Raster& r, const float* vert, const int* tri, int ntri, const float* b, float pix,
                 hipStream_t st, std::string& err) {
    const float w = b[2] - b[0], h = b[3] - b[1];
    // at most 4096 x 4096 pixels
    pix = std::max(pix, std::max(w, h) / 4095.0f);
    pix = std::max(pix, 1.0f);
    r.pix = pix;
    r.x0 = b[0];
    r.y0 = b[1];
    r.nx = (int)(w / pix) + 2;
    r.ny = (int)(h / pix) + 2;
    if (r.reserve(r.nx * r.ny, err)) return -3;
    int n = r.nx * r.ny;
    hipLaunchKernelGGL(k_raster_clear, dim3((n + 255) / 256), dim3(256), 0, st, r.z, n);
    if (ntri > 0)
        hipLaunchKernelGGL(k_raster_tris, dim3((ntri + 255) / 256), dim3(256), 0, st, vert, tri,
                           ntri, r.z, r.nx, r.ny, r.x0, r.y0, 1.0f / pix);
    HZ_HIP(hipGetLastError(), err);
    return 0;
}

void descriptors(const Raster& r, const CellFrame* cells, int ncell, uint16_t* desc,
                 hipStream_t st) {
    // ring k spans (R(k-1), R(k)] with R(k) = 500 m * 80^(k/7) and R(-1) =
    // kRingInner; its samples sit at R(k-1) * (R(k) / R(k-1))^(s / kSamples),
    // s = 1 .. kSamples
    DescTables tab;
    auto ring = [](int k) { return k < 0 ? (double)kRingInner : 500.0 * std::pow(80.0, k / 7.0); };
    for (int k = 0; k < kDescRings; k++)
        for (int s = 1; s <= kSamples; s++)
            tab.dist[k * kSamples + s - 1] =
                (float)(ring(k - 1) * std::pow(ring(k) / ring(k - 1), (double)s / kSamples));
    for (int j = 0; j < kDescAz; j++) {
        const double phi = j * (2.0 * 3.141592653589793 / kDescAz);
        tab.sin[j] = (float)std::sin(phi);
        tab.cos[j] = (float)std::cos(phi);
    }
    hipLaunchKernelGGL(k_descriptor, dim3((ncell + 127) / 128), dim3(128), 0, st, r.z, r.nx, r.ny,
                       r.x0, r.y0, 1.0f / r.pix, cells, ncell, tab, desc);
}

}  // namespace hz
