// C-ABI of the learned elevation bracket: the predictor object used by
// HZ_MODE_LEARNED and the native trainer that fits it on the device with the
// MFMA kernels of mlp.hip.  (No reference counterpart.)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/horayzon.h"
#include "hz_internal.h"
#include "learn_internal.h"
#include "mlp.h"

using namespace hz;

namespace hz {
int learn_fail(int code, const std::string& msg);  // defined in capi.cpp
}

#define LCHK(call)                                                                     \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return learn_fail(e_ == hipErrorOutOfMemory ? HZ_ENOMEM : HZ_EDEVICE,     \
                              std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

struct hz_trainer {
    hipStream_t st = nullptr;
    Mlp net;
    MlpAdam adam;
    MlpGrads g;
    MlpWork w;
    uint16_t* X = nullptr;
    float* T = nullptr;
    int n = 0;
    int* idx = nullptr;
    uint16_t* Xb = nullptr;
    float* Tb = nullptr;
    int batch_cap = 0;
    std::mt19937_64 rng;
};

namespace {

// parameter order of the exported (unpadded) layout
const size_t kNW1 = (size_t)kH1 * kIn, kNW2 = (size_t)kH2 * kH1, kNW3 = (size_t)kOut * kH2;

void split(const float* flat, const float** w1, const float** b1, const float** w2,
           const float** b2, const float** w3, const float** b3) {
    *w1 = flat;
    *b1 = *w1 + kNW1;
    *w2 = *b1 + kH1;
    *b2 = *w2 + kNW2;
    *w3 = *b2 + kH2;
    *b3 = *w3 + kNW3;
}

int trainer_batch(hz_trainer* t, int m) {
    std::string err;
    if (t->w.reserve(m, err)) return learn_fail(HZ_ENOMEM, err);
    if (m > t->batch_cap) {
        if (t->Xb) (void)hipFree(t->Xb);
        if (t->Tb) (void)hipFree(t->Tb);
        LCHK(hipMalloc(&t->Xb, (size_t)m * kIn * sizeof(uint16_t)));
        LCHK(hipMalloc(&t->Tb, (size_t)m * kOut * sizeof(float)));
        t->batch_cap = m;
    }
    return 0;
}

// gradients (device, padded layout) -> host, exported layout
int grads_to_host(hz_trainer* t, float* out) {
    std::vector<float> all(Mlp::num_floats());
    LCHK(hipMemcpyAsync(all.data(), t->g.base, all.size() * sizeof(float), hipMemcpyDeviceToHost,
                        t->st));
    LCHK(hipStreamSynchronize(t->st));
    float* o = out;
    std::memcpy(o, all.data(), kNW1 * 4);
    o += kNW1;
    std::memcpy(o, all.data() + Mlp::off_b1(), kH1 * 4);
    o += kH1;
    std::memcpy(o, all.data() + Mlp::off_w2(), kNW2 * 4);
    o += kNW2;
    std::memcpy(o, all.data() + Mlp::off_b2(), kH2 * 4);
    o += kH2;
    std::memcpy(o, all.data() + Mlp::off_w3(), kNW3 * 4);
    o += kNW3;
    std::memcpy(o, all.data() + Mlp::off_b3(), kOut * 4);
    return 0;
}

float host_loss(hz_trainer* t, int m) {
    int nb = (m * kOutPad + 255) / 256;
    std::vector<float> ls(nb);
    if (hipMemcpyAsync(ls.data(), t->w.lsum, nb * sizeof(float), hipMemcpyDeviceToHost, t->st) !=
        hipSuccess)
        return NAN;
    if (hipStreamSynchronize(t->st) != hipSuccess) return NAN;
    double s = 0.0;
    for (float v : ls) s += v;
    return (float)(s / ((double)m * kOut));
}

}  // namespace

extern "C" {

long long hz_predictor_num_params(void) {
    return (long long)(kNW1 + kH1 + kNW2 + kH2 + kNW3 + kOut);
}

int hz_predictor_create(const float* params, hz_predictor** out) {
    if (!out || !params) return learn_fail(HZ_EINVAL, "NULL argument");
    *out = nullptr;
    hz_predictor* p = new hz_predictor();
    std::string err;
    if (p->net.alloc(err)) {
        delete p;
        return learn_fail(HZ_ENOMEM, err);
    }
    const float *w1, *b1, *w2, *b2, *w3, *b3;
    split(params, &w1, &b1, &w2, &b2, &w3, &b3);
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        p->net.free_all();
        delete p;
        return learn_fail(HZ_EDEVICE, "hipStreamCreate failed");
    }
    int rc = p->net.upload(w1, b1, w2, b2, w3, b3, st, err);
    (void)hipStreamDestroy(st);
    if (rc) {
        p->net.free_all();
        delete p;
        return learn_fail(HZ_EDEVICE, err);
    }
    *out = p;
    return HZ_OK;
}

void hz_predictor_destroy(hz_predictor* p) {
    if (!p) return;
    p->net.free_all();
    delete p;
}

int hz_trainer_create(unsigned long long seed, hz_trainer** out) {
    if (!out) return learn_fail(HZ_EINVAL, "NULL argument");
    *out = nullptr;
    hz_trainer* t = new hz_trainer();
    std::string err;
    t->rng.seed(seed);
    if (hipStreamCreateWithFlags(&t->st, hipStreamNonBlocking) != hipSuccess) {
        delete t;
        return learn_fail(HZ_EDEVICE, "hipStreamCreate failed");
    }
    if (t->net.alloc(err)) {
        hz_trainer_destroy(t);
        return learn_fail(HZ_ENOMEM, err);
    }
    size_t nf = Mlp::num_floats();
    if (hipMalloc(&t->g.base, nf * 4) != hipSuccess || hipMalloc(&t->adam.m, nf * 4) != hipSuccess ||
        hipMalloc(&t->adam.v, nf * 4) != hipSuccess) {
        hz_trainer_destroy(t);
        return learn_fail(HZ_ENOMEM, "hipMalloc failed");
    }
    t->g.w1 = t->g.base;
    t->g.w2 = t->g.base + Mlp::off_w2();
    t->g.w3 = t->g.base + Mlp::off_w3();
    t->g.b1 = t->g.base + Mlp::off_b1();
    t->g.b2 = t->g.base + Mlp::off_b2();
    t->g.b3 = t->g.base + Mlp::off_b3();
    (void)hipMemsetAsync(t->adam.m, 0, nf * 4, t->st);
    (void)hipMemsetAsync(t->adam.v, 0, nf * 4, t->st);
    (void)hipMemsetAsync(t->g.base, 0, nf * 4, t->st);
    // He initialisation for the GELU layers, 1/sqrt(fan_in) for the output
    std::vector<float> p(hz_predictor_num_params(), 0.0f);
    float *w1 = p.data(), *b1 = w1 + kNW1, *w2 = b1 + kH1, *b2 = w2 + kNW2, *w3 = b2 + kH2,
          *b3 = w3 + kNW3;
    (void)b1;
    (void)b2;
    (void)b3;
    std::normal_distribution<float> nd(0.0f, 1.0f);
    for (size_t i = 0; i < kNW1; i++) w1[i] = nd(t->rng) * std::sqrt(2.0f / kIn);
    for (size_t i = 0; i < kNW2; i++) w2[i] = nd(t->rng) * std::sqrt(2.0f / kH1);
    for (size_t i = 0; i < kNW3; i++) w3[i] = nd(t->rng) * std::sqrt(1.0f / kH2);
    int rc = hz_trainer_set_weights(t, p.data());
    if (rc) {
        hz_trainer_destroy(t);
        return rc;
    }
    *out = t;
    return HZ_OK;
}

void hz_trainer_destroy(hz_trainer* t) {
    if (!t) return;
    if (t->st) (void)hipStreamSynchronize(t->st);
    t->net.free_all();
    t->w.free_all();
    void* b[] = {t->g.base, t->adam.m, t->adam.v, t->X, t->T, t->idx, t->Xb, t->Tb};
    for (void* q : b)
        if (q) (void)hipFree(q);
    if (t->st) (void)hipStreamDestroy(t->st);
    delete t;
}

int hz_trainer_set_weights(hz_trainer* t, const float* params) {
    if (!t || !params) return learn_fail(HZ_EINVAL, "NULL argument");
    const float *w1, *b1, *w2, *b2, *w3, *b3;
    split(params, &w1, &b1, &w2, &b2, &w3, &b3);
    std::string err;
    if (t->net.upload(w1, b1, w2, b2, w3, b3, t->st, err)) return learn_fail(HZ_EDEVICE, err);
    return HZ_OK;
}

int hz_trainer_get_weights(hz_trainer* t, float* params) {
    if (!t || !params) return learn_fail(HZ_EINVAL, "NULL argument");
    float* w1 = params;
    float *b1 = w1 + kNW1, *w2 = b1 + kH1, *b2 = w2 + kNW2, *w3 = b2 + kH2, *b3 = w3 + kNW3;
    std::string err;
    if (t->net.download(w1, b1, w2, b2, w3, b3, t->st, err)) return learn_fail(HZ_EDEVICE, err);
    return HZ_OK;
}

int hz_trainer_set_data(hz_trainer* t, const unsigned short* X, const float* T, int n) {
    if (!t || !X || !T || n < 1) return learn_fail(HZ_EINVAL, "bad training data");
    if (t->X) (void)hipFree(t->X);
    if (t->T) (void)hipFree(t->T);
    if (t->idx) (void)hipFree(t->idx);
    t->X = nullptr;
    t->T = nullptr;
    t->idx = nullptr;
    LCHK(hipMalloc(&t->X, (size_t)n * kIn * sizeof(uint16_t)));
    LCHK(hipMalloc(&t->T, (size_t)n * kOut * sizeof(float)));
    LCHK(hipMalloc(&t->idx, (size_t)n * sizeof(int)));
    LCHK(hipMemcpyAsync(t->X, X, (size_t)n * kIn * 2, hipMemcpyHostToDevice, t->st));
    LCHK(hipMemcpyAsync(t->T, T, (size_t)n * kOut * 4, hipMemcpyHostToDevice, t->st));
    LCHK(hipStreamSynchronize(t->st));
    t->n = n;
    return HZ_OK;
}

int hz_trainer_fit(hz_trainer* t, int epochs, int batch, float lr, float* epoch_loss) {
    if (!t || t->n < 1) return learn_fail(HZ_EINVAL, "no training data");
    if (epochs < 1 || batch < 64 || batch % 64 != 0)
        return learn_fail(HZ_EINVAL, "epochs >= 1 and batch a multiple of 64 are required");
    batch = std::min(batch, std::max(64, t->n / 64 * 64));
    int rc = trainer_batch(t, batch);
    if (rc) return rc;
    std::vector<int> perm(t->n);
    std::iota(perm.begin(), perm.end(), 0);
    const int nbatch = t->n / batch;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), t->rng);
        LCHK(hipMemcpyAsync(t->idx, perm.data(), (size_t)t->n * 4, hipMemcpyHostToDevice, t->st));
        // cosine decay of the step size over the run
        for (int b = 0; b < std::max(1, nbatch); b++) {
            const int m = std::min(batch, t->n - b * batch);
            if (m < 64) break;
            const int mm = m / 64 * 64;
            const float prog = ((float)ep * nbatch + b) / (float)(epochs * std::max(1, nbatch));
            const float lr_t = lr * 0.5f * (1.0f + std::cos(3.14159265f * prog));
            mlp_gather(t->X, t->T, t->idx + (size_t)b * batch, mm, t->Xb, t->Tb, t->st);
            mlp_forward(t->net, t->Xb, mm, t->w, t->st);
            mlp_backward(t->net, t->Xb, t->Tb, mm, t->w, t->g, t->st);
            mlp_adam(t->net, t->adam, t->g, lr_t, t->st);
        }
        LCHK(hipGetLastError());
        if (epoch_loss) epoch_loss[ep] = host_loss(t, std::min(batch, t->n / 64 * 64));
        else LCHK(hipStreamSynchronize(t->st));
    }
    return HZ_OK;
}

int hz_trainer_gradients(hz_trainer* t, const unsigned short* X, const float* T, int n,
                         float* grads, float* loss) {
    if (!t || !X || !T || !grads || n < 8 || n % 8 != 0)
        return learn_fail(HZ_EINVAL, "n must be a positive multiple of 8");
    int rc = trainer_batch(t, n);
    if (rc) return rc;
    LCHK(hipMemcpyAsync(t->Xb, X, (size_t)n * kIn * 2, hipMemcpyHostToDevice, t->st));
    LCHK(hipMemcpyAsync(t->Tb, T, (size_t)n * kOut * 4, hipMemcpyHostToDevice, t->st));
    mlp_forward(t->net, t->Xb, n, t->w, t->st);
    mlp_backward(t->net, t->Xb, t->Tb, n, t->w, t->g, t->st);
    LCHK(hipGetLastError());
    if (loss) *loss = host_loss(t, n);
    return grads_to_host(t, grads);
}

int hz_trainer_predict(hz_trainer* t, const unsigned short* X, int n, float* out) {
    if (!t || !X || !out || n < 1) return learn_fail(HZ_EINVAL, "bad arguments");
    const int chunk = 1 << 16;
    int rc = trainer_batch(t, std::min(n, chunk));
    if (rc) return rc;
    float* y = nullptr;
    LCHK(hipMalloc(&y, (size_t)std::min(n, chunk) * kOutPad * sizeof(float)));
    struct Free {
        float* p;
        ~Free() { (void)hipFree(p); }
    } free_y{y};  // also on the error returns below
    std::vector<float> tmp((size_t)std::min(n, chunk) * kOutPad);
    for (int b = 0; b < n; b += chunk) {
        int m = std::min(chunk, n - b);
        LCHK(hipMemcpyAsync(t->Xb, X + (size_t)b * kIn, (size_t)m * kIn * 2, hipMemcpyHostToDevice,
                            t->st));
        mlp_infer(t->net, t->Xb, m, y, t->st);
        LCHK(hipMemcpyAsync(tmp.data(), y, (size_t)m * kOutPad * 4, hipMemcpyDeviceToHost, t->st));
        LCHK(hipStreamSynchronize(t->st));
        for (int i = 0; i < m; i++)
            std::memcpy(out + (size_t)(b + i) * kOut, tmp.data() + (size_t)i * kOutPad, kOut * 4);
    }
    return HZ_OK;
}

}  // extern "C"
