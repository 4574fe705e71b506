// Predictor network of the learned elevation bracket (mlp.hip) and its
// descriptor (desc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace hz {

constexpr int kDescAz = 24;    // descriptor azimuths (one per output)
constexpr int kDescRings = 8;  // radial rings per azimuth
constexpr int kIn = kDescAz * kDescRings;  // 192
constexpr int kH1 = 256;
constexpr int kH2 = 256;
constexpr int kOut = 24;
constexpr int kOutPad = 32;  // output layer padded to two 16-wide MFMA tiles

// Weights: FP32 masters, BF16 copies (w*b, [out][in]) and BF16 transposes
// (w*t, [in][out]) for the backward pass.
struct Mlp {
    float* wf = nullptr;  // one allocation: w1 w2 w3 b1 b2 b3
    uint16_t* wb = nullptr;
    float *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
    uint16_t *w1b = nullptr, *w2b = nullptr, *w3b = nullptr;
    uint16_t *w1t = nullptr, *w2t = nullptr, *w3t = nullptr;
    static constexpr size_t off_w2() { return (size_t)kH1 * kIn; }
    static constexpr size_t off_w3() { return off_w2() + (size_t)kH2 * kH1; }
    static constexpr size_t off_b1() { return off_w3() + (size_t)kOutPad * kH2; }
    static constexpr size_t off_b2() { return off_b1() + kH1; }
    static constexpr size_t off_b3() { return off_b2() + kH2; }
    static constexpr size_t num_floats() { return off_b3() + kOutPad; }
    int alloc(std::string& err);
    void free_all();
    // host arrays: w1 [256][192], b1 [256], w2 [256][256], b2 [256], w3 [24][256], b3 [24]
    int upload(const float* w1, const float* b1, const float* w2, const float* b2, const float* w3,
               const float* b3, hipStream_t st, std::string& err);
    int download(float* w1, float* b1, float* w2, float* b2, float* w3, float* b3, hipStream_t st,
                 std::string& err);
    void refresh_bf16(hipStream_t st);
};

struct MlpGrads {  // same layout as Mlp::wf
    float* base = nullptr;
    float *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *b1 = nullptr, *b2 = nullptr, *b3 = nullptr;
};

struct MlpAdam {
    float* m = nullptr;
    float* v = nullptr;
    int t = 0;
};

struct MlpWork {
    int cap = 0;
    float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *dh = nullptr;
    uint16_t *h1 = nullptr, *h2 = nullptr, *dy = nullptr, *dz = nullptr, *tr_a = nullptr,
             *tr_b = nullptr;
    float* lsum = nullptr;
    int reserve(int m, std::string& err);
    void free_all();
};

void mlp_gemm(const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N, int K,
              const float* bias, int act, float* C, int ldc, float* Z, uint16_t* H, int ldh,
              hipStream_t st);
void mlp_transpose(const uint16_t* in, int rows, int cols, int ldi, uint16_t* out, int ldo,
                   hipStream_t st);
void mlp_forward(const Mlp& net, const uint16_t* X, int M, MlpWork& w, hipStream_t st);
void mlp_infer(const Mlp& net, const uint16_t* X, int M, float* out, hipStream_t st);
void mlp_backward(const Mlp& net, const uint16_t* X, const float* T, int M, MlpWork& w,
                  MlpGrads& g, hipStream_t st);
void mlp_adam(Mlp& net, MlpAdam& opt, const MlpGrads& g, float lr, hipStream_t st);
void mlp_gather(const uint16_t* X, const float* T, const int* idx, int n, uint16_t* Xb, float* Tb,
                hipStream_t st);

// ---------------------------------------------------------------------------
// Descriptor (desc.hip)
// ---------------------------------------------------------------------------

// Max-height raster of the FP32 scene in the ENU x/y plane.
struct Raster {
    unsigned* z = nullptr;  // order-preserving encoded FP32 heights
    int nx = 0, ny = 0, cap = 0;
    float x0 = 0, y0 = 0, pix = 0;
    int reserve(int n, std::string& err);
    void free_all();
};

// Inner radius of the descriptor's first ring [m]; ring k (0 .. 7) ends at
// 500 m * 80^(k/7), the last at 40 km.
constexpr float kRingInner = 250.0f;

int raster_build(Raster& r, const float* vert, const int* tri, int ntri, const float* bounds_host,
                 float pix, hipStream_t st, std::string& err);
// desc [ncell][kIn] BF16: for azimuth j (clockwise from north, 15 deg apart)
// and ring k, the largest elevation angle (deg / 10) of the raster samples
// in that ring seen from the cell's ray origin.
void descriptors(const Raster& r, const struct CellFrame* cells, int ncell, uint16_t* desc,
                 hipStream_t st);

}  // namespace hz
