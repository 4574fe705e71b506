// Calibration of the memory path that the walk's node fetches use, for the
// roofline of the trace kernel: dependent random reads of 64 bytes (four
// 16-byte loads per lane) from an array of 128-byte records, the way a
// traversal step reads a node's boxes and references.  Each lane follows
// its own chain (the next record's index comes from the bits just read), at
// the trace kernel's occupancy (four wavefronts per SIMD, set by the LDS
// size) and lane activity (about 60% of the lanes per step), with a short
// FP32 computation per step.  The rate does not depend on the computation
// (0 to 200 FMAs per step) nor on the occupancy from two to eight
// wavefronts per SIMD: it is the GPU's rate for such gathers.

#include "hz_internal.h"

namespace hz {
namespace {

constexpr int kCalibBlock = 128;
constexpr unsigned kActive = 614;  // of 1024: the share of lanes that step

__device__ __forceinline__ unsigned mix32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void k_calib_fill(uint4* __restrict__ rec, long long n16, unsigned seed) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n16;
         i += (long long)gridDim.x * blockDim.x) {
        const unsigned h = mix32((unsigned)i ^ seed);
        rec[i] = make_uint4(mix32(h), mix32(h + 1u), mix32(h + 2u), mix32(h + 3u));
    }
}

// hot: the first `hot` records take every other step (a resident set), the
// rest spread over all n records.
__global__ __launch_bounds__(kCalibBlock) void k_calib_chase(const float4* __restrict__ rec,
                                                             unsigned n, unsigned hot, int iters,
                                                             float* __restrict__ sink) {
    extern __shared__ int pad[];  // sized by the launch: sets the occupancy
    const unsigned gid = blockIdx.x * kCalibBlock + threadIdx.x;
    unsigned idx = mix32(gid) % n;
    float acc = 0.f;
    if (threadIdx.x == 0) pad[0] = 0;
    for (int it = 0; it < iters; it++) {
        if ((mix32(gid * 977u + (unsigned)it) & 1023u) < kActive) {
            const float4* p = rec + (size_t)idx * 8;
            const float4 a = p[0], b = p[1], c = p[2], d = p[3];
            float x = a.y + b.y * c.z - d.w;
#pragma unroll
            for (int q = 0; q < 100; q++) x = __builtin_fmaf(x, 1.0001f, 0.5f);
            acc += x;
            const unsigned r = __float_as_uint(a.x) ^ mix32(idx + (unsigned)it);
            idx = (r & 1u) ? (r >> 1) % hot : (r >> 1) % n;
        }
    }
    sink[gid] = acc;
}

}  // namespace

int gather_chase_rate(long long num_nodes, long long hot_nodes, int iters, double* rate,
                      std::string& err) {
    if (num_nodes < 1024 || num_nodes > (1ll << 27) || hot_nodes < 1 || hot_nodes > num_nodes ||
        iters < 1 || iters > 100000 || !rate) {
        err = "hz_calibrate_gather: num_nodes in [1024, 2^27], hot_nodes in [1, num_nodes], "
              "iters in [1, 100000]";
        return -1;
    }
    int dev = 0, cus = 0, lds_cu = 0;
    HZ_HIP(hipGetDevice(&dev), err);
    HZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), err);
    HZ_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev),
           err);
    // four wavefronts per SIMD: eight 128-lane workgroups per CU, each with
    // an eighth of the CU's LDS
    const int blocks_cu = 8;
    const size_t lds = std::max<size_t>(1024, ((size_t)lds_cu / blocks_cu) & ~(size_t)1023);
    const int grid = std::max(1, cus) * blocks_cu;
    float4* rec = nullptr;
    float* sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = [&]() {
        if (rec) (void)hipFree(rec);
        if (sink) (void)hipFree(sink);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    auto run = [&]() -> int {
        HZ_HIP(hipMalloc(&rec, (size_t)num_nodes * 128), err);
        HZ_HIP(hipMalloc(&sink, (size_t)grid * kCalibBlock * sizeof(float)), err);
        HZ_HIP(hipEventCreate(&e0), err);
        HZ_HIP(hipEventCreate(&e1), err);
        hipLaunchKernelGGL(k_calib_fill, dim3(1024), dim3(256), 0, 0,
                           reinterpret_cast<uint4*>(rec), num_nodes * 8, 12345u);
        HZ_HIP(hipGetLastError(), err);
        const unsigned n = (unsigned)num_nodes, hot = (unsigned)hot_nodes;
        hipLaunchKernelGGL(k_calib_chase, dim3(grid), dim3(kCalibBlock), lds, 0, rec, n, hot,
                           iters, sink);  // warm-up
        HZ_HIP(hipGetLastError(), err);
        constexpr int kTimed = 3;
        HZ_HIP(hipEventRecord(e0, 0), err);
        for (int k = 0; k < kTimed; k++)
            hipLaunchKernelGGL(k_calib_chase, dim3(grid), dim3(kCalibBlock), lds, 0, rec, n, hot,
                               iters, sink);
        HZ_HIP(hipEventRecord(e1, 0), err);
        HZ_HIP(hipEventSynchronize(e1), err);
        HZ_HIP(hipGetLastError(), err);
        float ms = 0.f;
        HZ_HIP(hipEventElapsedTime(&ms, e0, e1), err);
        const double steps =
            (double)grid * kCalibBlock * iters * ((double)kActive / 1024.0) * kTimed;
        *rate = ms > 0.f ? steps / (ms * 1e-3) : 0.0;
        return 0;
    };
    const int rc = run();
    release();
    return rc;
}

}  // namespace hz
