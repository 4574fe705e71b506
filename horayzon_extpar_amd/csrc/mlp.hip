// Elevation-bracket predictor: a 192 -> 256 -> 256 -> 24 MLP with GELU,
// BF16 weights and activations, FP32 accumulation, trained and evaluated by
// the kernels of this file.  (No reference counterpart: the reference always
// runs the full elevation search, mo_lradtopo_horayzon.cpp:490-555.)
//
// GEMM.  One kernel computes C[M][N] = A[M][K] * B[N][K]^T with both
// operands K-contiguous in BF16, on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16).  A wavefront owns a 32 x 32 tile of C (2 x 2
// MFMA tiles) and walks K in steps of 32; every lane loads its 8 consecutive
// K elements of an A row / B row as one 16-byte load, which is the operand
// layout of the instruction:
//   A: lane l holds A[row = l % 16][k = 8 (l / 16) .. +7]
//   B: lane l holds B[col = l % 16][k = 8 (l / 16) .. +7]
//   D: lane l holds D[row = 4 (l / 16) + i][col = l % 16], i = 0..3
// The epilogue adds a bias, applies GELU and writes any of: the FP32 result,
// the FP32 pre-activation and the BF16 activation.  The backward GEMMs use
// the same kernel on transposed copies (transpose kernel below).  Inference
// runs all three layers in one kernel (k_mlp_fused) with the same MFMA
// chains, so its predictions equal those of the three GEMMs.

#include <cstring>

#include "hz_internal.h"
#include "mlp.h"

namespace hz {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even (inputs here are finite)
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p, bool ok) {
    if (!ok) {
        bf16x8 z;
        for (int i = 0; i < 8; i++) z[i] = (__bf16)0.0f;
        return z;
    }
    uint4 v = *reinterpret_cast<const uint4*>(p);
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

// 4 waves per workgroup, 64 x 64 C tile per workgroup.
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ B, int ldb,
                                                 int M, int N, int K,
                                                 const float* __restrict__ bias, int act,
                                                 float* __restrict__ C, int ldc,
                                                 float* __restrict__ Z, uint16_t* __restrict__ H,
                                                 int ldh) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int r = lane & 15, g = lane >> 4;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        const int k = k0 + 8 * g;
        const bool kok = k < K;  // K is a multiple of 8
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; i++) {
            const int row = m0 + 16 * i + r;
            a[i] = load8(A + (size_t)row * lda + k, kok && row < M);
        }
        for (int j = 0; j < 2; j++) {
            const int col = n0 + 16 * j + r;
            b[j] = load8(B + (size_t)col * ldb + k, kok && col < N);
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            const int col = n0 + 16 * j + r;
            if (col >= N) continue;
            const float bb = bias ? bias[col] : 0.0f;
            for (int q = 0; q < 4; q++) {
                const int row = m0 + 16 * i + 4 * g + q;
                if (row >= M) continue;
                const float z = acc[i][j][q] + bb;
                if (C) C[(size_t)row * ldc + col] = act ? gelu(z) : z;
                if (Z) Z[(size_t)row * ldc + col] = z;
                if (H) H[(size_t)row * ldh + col] = f2bf(act ? gelu(z) : z);
            }
        }
}

// out[c][r] = in[r][c] (BF16), rows x cols
__global__ void k_transpose(const uint16_t* __restrict__ in, int rows, int cols, int ldi,
                            uint16_t* __restrict__ out, int ldo) {
    __shared__ uint16_t t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int y = threadIdx.y; y < 32; y += 8) {
        int rr = r0 + y, cc = c0 + threadIdx.x;
        t[y][threadIdx.x] = (rr < rows && cc < cols) ? in[(size_t)rr * ldi + cc] : 0;
    }
    __syncthreads();
    for (int y = threadIdx.y; y < 32; y += 8) {
        int cc = c0 + y, rr = r0 + threadIdx.x;
        if (cc < cols && rr < rows) out[(size_t)cc * ldo + rr] = t[threadIdx.x][y];
    }
}

// loss = mean over rows and the first nout columns of (Y - T)^2;
// dY = d loss / dY as BF16 (zero in the padding columns); partial sums of
// the loss per block in lsum.
__global__ void k_mse(const float* __restrict__ Y, int ldy, const float* __restrict__ T,
                      int nout, int npad, int M, uint16_t* __restrict__ dY,
                      float* __restrict__ lsum) {
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    const float scale = 2.0f / ((float)M * (float)nout);
    if (i < M * npad) {
        const int row = i / npad, col = i % npad;
        float d = 0.0f;
        if (col < nout) {
            d = Y[(size_t)row * ldy + col] - T[(size_t)row * nout + col];
            l = d * d;
        }
        dY[i] = f2bf(scale * d);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lsum[blockIdx.x] = red[0];
}

// dZ = dH * gelu'(Z) as BF16
__global__ void k_gelu_back(const float* __restrict__ dH, const float* __restrict__ Z, size_t n,
                            uint16_t* __restrict__ dZ) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dZ[i] = f2bf(dH[i] * gelu_grad(Z[i]));
}

// column sums of a BF16 matrix (bias gradients): out[c] = sum_r in[r][c]
__global__ void k_colsum(const uint16_t* __restrict__ in, int rows, int cols,
                         float* __restrict__ out) {
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int part = threadIdx.x >> 6;  // 4 partial sums
    __shared__ float red[4][64];
    float s = 0.0f;
    if (c < cols)
        for (int r = part; r < rows; r += 4) s += bf2f(in[(size_t)r * cols + c]);
    red[part][threadIdx.x & 63] = s;
    __syncthreads();
    if (part == 0 && c < cols)
        out[c] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Adam on FP32 master weights; refreshes the BF16 copy [rows][cols] and its
// transpose.  grad layout = weight layout.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int rows, int cols, float lr, float b1,
                       float b2, float eps, float c1, float c2, uint16_t* __restrict__ wb,
                       uint16_t* __restrict__ wbt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * cols) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
    if (wbt) wbt[(size_t)(i % cols) * rows + i / cols] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ in, size_t n, uint16_t* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// rows of a dataset (BF16 features, FP32 targets) gathered by index
__global__ void k_gather(const uint16_t* __restrict__ X, const float* __restrict__ T,
                         const int* __restrict__ idx, int n, int kin, int nout,
                         uint16_t* __restrict__ Xb, float* __restrict__ Tb) {
    const int row = blockIdx.x;
    if (row >= n) return;
    const size_t src = (size_t)idx[row];
    for (int c = threadIdx.x; c < kin; c += blockDim.x) Xb[(size_t)row * kin + c] = X[src * kin + c];
    for (int c = threadIdx.x; c < nout; c += blockDim.x) Tb[(size_t)row * nout + c] = T[src * nout + c];
}

inline dim3 gemm_grid(int M, int N) { return dim3((N + 63) / 64, (M + 63) / 64); }

// Fused inference.  One workgroup (4 wavefronts) carries 64 rows through all
// three layers; the rows' activations stay in one LDS buffer (X, then H1,
// then H2, each overwritten once every wavefront has read it) and the BF16
// weights (240 KB, resident in L2) are read by every workgroup.  Wavefront
// w computes columns 64 w .. 64 w + 63 of the hidden layers (4 x 4 MFMA
// tiles over all 64 rows) and rows 16 w .. 16 w + 15 of the output layer.
// The MFMA chain per tile walks K in the order of k_gemm_nt, so the results
// equal the three-GEMM path bit for bit.
constexpr int kFRows = 64;
constexpr int kFLd = kH1 + 8;  // LDS row stride (elements): +16 bytes against bank conflicts
static_assert(kH1 == kH2 && kIn <= kH1, "one LDS buffer holds X, H1 and H2");

template <int K>
__device__ __forceinline__ void fused_tile(const uint16_t* s, const uint16_t* __restrict__ W,
                                           int n0, int r, int g, f32x4 (&acc)[4][4]) {
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a[4], b[4];
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = load8(W + (size_t)(n0 + 16 * j + r) * K + k0 + 8 * g, true);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 v = *reinterpret_cast<const uint4*>(s + (16 * i + r) * kFLd + k0 + 8 * g);
            __builtin_memcpy(&a[i], &v, 16);
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
}

__device__ __forceinline__ void fused_store(uint16_t* s, const float* __restrict__ bias, int n0,
                                            int r, int g, const f32x4 (&acc)[4][4]) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int col = n0 + 16 * j + r;
        const float bb = bias[col];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) s[(16 * i + 4 * g + q) * kFLd + col] = f2bf(gelu(acc[i][j][q] + bb));
    }
}

__global__ __launch_bounds__(256) void k_mlp_fused(const uint16_t* __restrict__ X, int M,
                                                   const uint16_t* __restrict__ w1,
                                                   const float* __restrict__ b1,
                                                   const uint16_t* __restrict__ w2,
                                                   const float* __restrict__ b2,
                                                   const uint16_t* __restrict__ w3,
                                                   const float* __restrict__ b3,
                                                   float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) uint16_t s[kFRows * kFLd];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int row0 = blockIdx.x * kFRows;
    // X rows (zero beyond M), 16 bytes per thread and step
    for (int i = threadIdx.x; i < kFRows * (kIn / 8); i += 256) {
        const int row = i / (kIn / 8), c8 = i % (kIn / 8);
        uint4 v = {0u, 0u, 0u, 0u};
        if (row0 + row < M) v = *reinterpret_cast<const uint4*>(X + (size_t)(row0 + row) * kIn + 8 * c8);
        *reinterpret_cast<uint4*>(s + row * kFLd + 8 * c8) = v;
    }
    __syncthreads();
    const int n0 = 64 * wave;
    f32x4 acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    fused_tile<kIn>(s, w1, n0, r, g, acc);
    __syncthreads();  // every wavefront has read X
    fused_store(s, b1, n0, r, g, acc);
    __syncthreads();
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    fused_tile<kH1>(s, w2, n0, r, g, acc);
    __syncthreads();  // every wavefront has read H1
    fused_store(s, b2, n0, r, g, acc);
    __syncthreads();
    // output layer: 16 rows x 32 columns per wavefront
    f32x4 o[2] = {(f32x4){0.f, 0.f, 0.f, 0.f}, (f32x4){0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int k0 = 0; k0 < kH2; k0 += 32) {
        bf16x8 a, b[2];
        uint4 v = *reinterpret_cast<const uint4*>(s + (16 * wave + r) * kFLd + k0 + 8 * g);
        __builtin_memcpy(&a, &v, 16);
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = load8(w3 + (size_t)(16 * j + r) * kH2 + k0 + 8 * g, true);
#pragma unroll
        for (int j = 0; j < 2; j++) o[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], o[j], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int col = 16 * j + r;
        const float bb = b3[col];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int row = row0 + 16 * wave + 4 * g + q;
            if (row < M) out[(size_t)row * kOutPad + col] = o[j][q] + bb;
        }
    }
}
}  // namespace

void mlp_gemm(const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N, int K,
              const float* bias, int act, float* C, int ldc, float* Z, uint16_t* H, int ldh,
              hipStream_t st) {
    hipLaunchKernelGGL(k_gemm_nt, gemm_grid(M, N), dim3(256), 0, st, A, lda, B, ldb, M, N, K,
                       bias, act, C, ldc, Z, H, ldh);
}

void mlp_transpose(const uint16_t* in, int rows, int cols, int ldi, uint16_t* out, int ldo,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_transpose, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(32, 8), 0, st,
                       in, rows, cols, ldi, out, ldo);
}

// ---------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------

int Mlp::alloc(std::string& err) {
    const size_t nw1 = (size_t)kH1 * kIn, nw2 = (size_t)kH2 * kH1, nw3 = (size_t)kOutPad * kH2;
    size_t nf = nw1 + nw2 + nw3 + kH1 + kH2 + kOutPad;
    HZ_HIP(hipMalloc(&wf, nf * sizeof(float)), err);
    HZ_HIP(hipMalloc(&wb, (nw1 + nw2 + nw3) * 2 * sizeof(uint16_t)), err);
    w1 = wf;
    w2 = w1 + nw1;
    w3 = w2 + nw2;
    b1 = w3 + nw3;
    b2 = b1 + kH1;
    b3 = b2 + kH2;
    w1b = wb;
    w2b = w1b + nw1;
    w3b = w2b + nw2;
    w1t = w3b + nw3;
    w2t = w1t + nw1;
    w3t = w2t + nw2;
    return 0;
}

void Mlp::free_all() {
    if (wf) (void)hipFree(wf);
    if (wb) (void)hipFree(wb);
    wf = nullptr;
    wb = nullptr;
}

int Mlp::upload(const float* hw1, const float* hb1, const float* hw2, const float* hb2,
                const float* hw3, const float* hb3, hipStream_t st, std::string& err) {
    std::vector<float> w3p((size_t)kOutPad * kH2, 0.0f), b3p(kOutPad, 0.0f);
    std::memcpy(w3p.data(), hw3, sizeof(float) * kOut * kH2);
    std::memcpy(b3p.data(), hb3, sizeof(float) * kOut);
    HZ_HIP(hipMemcpyAsync(w1, hw1, sizeof(float) * kH1 * kIn, hipMemcpyHostToDevice, st), err);
    HZ_HIP(hipMemcpyAsync(w2, hw2, sizeof(float) * kH2 * kH1, hipMemcpyHostToDevice, st), err);
    HZ_HIP(hipMemcpyAsync(w3, w3p.data(), sizeof(float) * kOutPad * kH2, hipMemcpyHostToDevice, st), err);
    HZ_HIP(hipMemcpyAsync(b1, hb1, sizeof(float) * kH1, hipMemcpyHostToDevice, st), err);
    HZ_HIP(hipMemcpyAsync(b2, hb2, sizeof(float) * kH2, hipMemcpyHostToDevice, st), err);
    HZ_HIP(hipMemcpyAsync(b3, b3p.data(), sizeof(float) * kOutPad, hipMemcpyHostToDevice, st), err);
    refresh_bf16(st);
    HZ_HIP(hipStreamSynchronize(st), err);
    return 0;
}

void Mlp::refresh_bf16(hipStream_t st) {
    const size_t n[3] = {(size_t)kH1 * kIn, (size_t)kH2 * kH1, (size_t)kOutPad * kH2};
    const float* src[3] = {w1, w2, w3};
    uint16_t* dst[3] = {w1b, w2b, w3b};
    uint16_t* dstt[3] = {w1t, w2t, w3t};
    const int rows[3] = {kH1, kH2, kOutPad}, cols[3] = {kIn, kH1, kH2};
    for (int i = 0; i < 3; i++) {
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n[i] + 255) / 256)), dim3(256), 0, st, src[i],
                           n[i], dst[i]);
        mlp_transpose(dst[i], rows[i], cols[i], cols[i], dstt[i], rows[i], st);
    }
}

int Mlp::download(float* hw1, float* hb1, float* hw2, float* hb2, float* hw3, float* hb3,
                  hipStream_t st, std::string& err) {
    std::vector<float> w3p((size_t)kOutPad * kH2), b3p(kOutPad);
    HZ_HIP(hipMemcpyAsync(hw1, w1, sizeof(float) * kH1 * kIn, hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipMemcpyAsync(hw2, w2, sizeof(float) * kH2 * kH1, hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipMemcpyAsync(w3p.data(), w3, sizeof(float) * kOutPad * kH2, hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipMemcpyAsync(hb1, b1, sizeof(float) * kH1, hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipMemcpyAsync(hb2, b2, sizeof(float) * kH2, hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipMemcpyAsync(b3p.data(), b3, sizeof(float) * kOutPad, hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipStreamSynchronize(st), err);
    std::memcpy(hw3, w3p.data(), sizeof(float) * kOut * kH2);
    std::memcpy(hb3, b3p.data(), sizeof(float) * kOut);
    return 0;
}

int MlpWork::reserve(int m, std::string& err) {
    if (m <= cap) return 0;
    free_all();
    size_t M = (size_t)m;
    HZ_HIP(hipMalloc(&z1, M * kH1 * sizeof(float)), err);
    HZ_HIP(hipMalloc(&z2, M * kH2 * sizeof(float)), err);
    HZ_HIP(hipMalloc(&y, M * kOutPad * sizeof(float)), err);
    HZ_HIP(hipMalloc(&dh, M * kH1 * sizeof(float)), err);
    HZ_HIP(hipMalloc(&h1, M * kH1 * sizeof(uint16_t)), err);
    HZ_HIP(hipMalloc(&h2, M * kH2 * sizeof(uint16_t)), err);
    HZ_HIP(hipMalloc(&dy, M * kOutPad * sizeof(uint16_t)), err);
    HZ_HIP(hipMalloc(&dz, M * kH1 * sizeof(uint16_t)), err);
    HZ_HIP(hipMalloc(&tr_a, M * kH1 * sizeof(uint16_t)), err);
    HZ_HIP(hipMalloc(&tr_b, M * kH1 * sizeof(uint16_t)), err);
    HZ_HIP(hipMalloc(&lsum, ((M * kOutPad + 255) / 256 + 1) * sizeof(float)), err);
    cap = m;
    return 0;
}

void MlpWork::free_all() {
    void* p[] = {z1, z2, y, dh, h1, h2, dy, dz, tr_a, tr_b, lsum};
    for (void* q : p)
        if (q) (void)hipFree(q);
    *this = MlpWork();
}

// Forward pass of M rows of BF16 features X [M][kIn]; FP32 outputs in w.y
// ([M][kOutPad], first kOut columns used).  Keeps Z1, Z2, H1, H2 for the
// backward pass.
void mlp_forward(const Mlp& net, const uint16_t* X, int M, MlpWork& w, hipStream_t st) {
    mlp_gemm(X, kIn, net.w1b, kIn, M, kH1, kIn, net.b1, 1, nullptr, kH1, w.z1, w.h1, kH1, st);
    mlp_gemm(w.h1, kH1, net.w2b, kH1, M, kH2, kH1, net.b2, 1, nullptr, kH2, w.z2, w.h2, kH2, st);
    mlp_gemm(w.h2, kH2, net.w3b, kH2, M, kOutPad, kH2, net.b3, 0, w.y, kOutPad, nullptr, nullptr,
             0, st);
}

// Inference only (one fused launch, no intermediates in memory):
// predictions [M][kOutPad] in out.
void mlp_infer(const Mlp& net, const uint16_t* X, int M, float* out, hipStream_t st) {
    if (M > 0)
        hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)((M + kFRows - 1) / kFRows)), dim3(256), 0,
                           st, X, M, net.w1b, net.b1, net.w2b, net.b2, net.w3b, net.b3, out);
}

// Backward pass of the mean squared error against T [M][kOut]; gradients
// (FP32, same layouts as the weights) into g.  Requires mlp_forward first.
void mlp_backward(const Mlp& net, const uint16_t* X, const float* T, int M, MlpWork& w,
                  MlpGrads& g, hipStream_t st) {
    const int nb = (M * kOutPad + 255) / 256;
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, w.y, kOutPad, T, kOut, kOutPad, M, w.dy,
                       w.lsum);
    // layer 3: dW3 = dY^T H2, db3 = colsum dY, dH2 = dY W3
    mlp_transpose(w.dy, M, kOutPad, kOutPad, w.tr_a, M, st);  // [kOutPad][M]
    mlp_transpose(w.h2, M, kH2, kH2, w.tr_b, M, st);          // [kH2][M]
    mlp_gemm(w.tr_a, M, w.tr_b, M, kOutPad, kH2, M, nullptr, 0, g.w3, kH2, nullptr, nullptr, 0, st);
    hipLaunchKernelGGL(k_colsum, dim3((kOutPad + 63) / 64), dim3(256), 0, st, w.dy, M, kOutPad, g.b3);
    mlp_gemm(w.dy, kOutPad, net.w3t, kOutPad, M, kH2, kOutPad, nullptr, 0, w.dh, kH2, nullptr,
             nullptr, 0, st);
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)(((size_t)M * kH2 + 255) / 256)), dim3(256), 0, st,
                       w.dh, w.z2, (size_t)M * kH2, w.dz);
    // layer 2
    mlp_transpose(w.dz, M, kH2, kH2, w.tr_a, M, st);
    mlp_transpose(w.h1, M, kH1, kH1, w.tr_b, M, st);
    mlp_gemm(w.tr_a, M, w.tr_b, M, kH2, kH1, M, nullptr, 0, g.w2, kH1, nullptr, nullptr, 0, st);
    hipLaunchKernelGGL(k_colsum, dim3((kH2 + 63) / 64), dim3(256), 0, st, w.dz, M, kH2, g.b2);
    mlp_gemm(w.dz, kH2, net.w2t, kH2, M, kH1, kH2, nullptr, 0, w.dh, kH1, nullptr, nullptr, 0, st);
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)(((size_t)M * kH1 + 255) / 256)), dim3(256), 0, st,
                       w.dh, w.z1, (size_t)M * kH1, w.dz);
    // layer 1
    mlp_transpose(w.dz, M, kH1, kH1, w.tr_a, M, st);
    mlp_transpose(X, M, kIn, kIn, w.tr_b, M, st);
    mlp_gemm(w.tr_a, M, w.tr_b, M, kH1, kIn, M, nullptr, 0, g.w1, kIn, nullptr, nullptr, 0, st);
    hipLaunchKernelGGL(k_colsum, dim3((kH1 + 63) / 64), dim3(256), 0, st, w.dz, M, kH1, g.b1);
}

void mlp_adam(Mlp& net, MlpAdam& opt, const MlpGrads& g, float lr, hipStream_t st) {
    opt.t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)opt.t), c2 = 1.0f - powf(b2, (float)opt.t);
    struct P {
        float* w;
        float* m;
        float* v;
        const float* g;
        int rows, cols;
        uint16_t* wb;
        uint16_t* wbt;
    } ps[6] = {{net.w1, opt.m, opt.v, g.w1, kH1, kIn, net.w1b, net.w1t},
               {net.w2, opt.m + Mlp::off_w2(), opt.v + Mlp::off_w2(), g.w2, kH2, kH1, net.w2b, net.w2t},
               {net.w3, opt.m + Mlp::off_w3(), opt.v + Mlp::off_w3(), g.w3, kOutPad, kH2, net.w3b,
                net.w3t},
               {net.b1, opt.m + Mlp::off_b1(), opt.v + Mlp::off_b1(), g.b1, 1, kH1, nullptr, nullptr},
               {net.b2, opt.m + Mlp::off_b2(), opt.v + Mlp::off_b2(), g.b2, 1, kH2, nullptr, nullptr},
               {net.b3, opt.m + Mlp::off_b3(), opt.v + Mlp::off_b3(), g.b3, 1, kOutPad, nullptr,
                nullptr}};
    for (auto& p : ps) {
        int n = p.rows * p.cols;
        hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, p.w, p.m, p.v, p.g,
                           p.rows, p.cols, lr, b1, b2, eps, c1, c2, p.wb, p.wbt);
    }
}

void mlp_gather(const uint16_t* X, const float* T, const int* idx, int n, uint16_t* Xb, float* Tb,
                hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3(n), dim3(64), 0, st, X, T, idx, n, kIn, kOut, Xb, Tb);
}

}  // namespace hz
