// Large host <-> device copies between pageable caller memory and the device
// through three 16 MB slots of pinned memory: asynchronous copies move whole
// slots on the stream while items of the library's worker pool (up to 8 per
// slot) copy between the caller's memory and the other slots.  Pinned copies
// run near the PCIe rate and the host side is parallel (it also takes the
// first-touch page faults of a freshly allocated destination in parallel),
// where a plain copy from / to pageable memory goes through one runtime
// staging buffer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hz_internal.h"

namespace hz {

namespace {
constexpr size_t kChunk = (size_t)16 << 20;
constexpr size_t kDirect = (size_t)8 << 20;  // below this: one plain copy
}  // namespace

void Stager::release() {
    if (buf) (void)hipHostFree(buf);
    buf = nullptr;
    for (auto& e : ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

hipError_t Stager::copy(void* dst, const void* src, size_t bytes, bool to_device,
                        hipStream_t st) {
    const hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    auto plain = [&]() {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
        return e != hipSuccess ? e : hipStreamSynchronize(st);
    };
    if (bytes < kDirect) return plain();
    hipError_t e = hipSuccess;
    if (!buf) {
        // pinned memory is a limited resource: without it, copy directly
        if (hipHostMalloc(&buf, kChunk * kSlots) != hipSuccess) {
            buf = nullptr;
            (void)hipGetLastError();
            return plain();
        }
        for (auto& x : ev) {
            e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
    }
    char* slot = static_cast<char*>(buf);
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    const int nch = (int)((bytes + kChunk - 1) / kChunk);
    auto len = [&](int k) { return std::min(kChunk, bytes - (size_t)k * kChunk); };
    auto sl = [&](int k) { return slot + (size_t)(k % kSlots) * kChunk; };
    // device side of chunk k (slot <-> device), followed by the slot's event
    auto issue = [&](int k) {
        hipError_t r = to_device ? hipMemcpyAsync(d + (size_t)k * kChunk, sl(k), len(k), kind, st)
                                 : hipMemcpyAsync(sl(k), s + (size_t)k * kChunk, len(k), kind, st);
        return r != hipSuccess ? r : hipEventRecord(ev[k % kSlots], st);
    };
    // host side of chunk k: nt slices, each a pool item that never waits
    const int nt = std::max(1, std::min(pool_threads(), 8));
    std::unique_ptr<TaskGroup[]> grp(new TaskGroup[nch]);
    int submitted = 0;
    auto host = [&](int k) {
        const size_t n = len(k);
        char* hd = to_device ? sl(k) : d + (size_t)k * kChunk;
        const char* hs = to_device ? s + (size_t)k * kChunk : sl(k);
        for (int t = 0; t < nt; t++) {
            const size_t b = n / nt * t, en = t == nt - 1 ? n : n / nt * (t + 1);
            if (b < en) pool_submit(grp[k], [hd, hs, b, en]() { std::memcpy(hd + b, hs + b, en - b); });
        }
        submitted = k + 1;
    };
    if (to_device) {
        // host fills slot k once the device copy of chunk k - kSlots out of it
        // is over (up to kSlots chunks ahead); the device copies chunk k once
        // its host part is done
        int next = 0;
        for (int k = 0; k < nch && e == hipSuccess; k++) {
            while (next < nch && next < k + kSlots && e == hipSuccess) {
                if (next >= kSlots) e = hipEventSynchronize(ev[next % kSlots]);
                if (e == hipSuccess) host(next++);
            }
            if (e != hipSuccess) break;
            pool_wait(grp[k]);
            e = issue(k);
        }
    } else {
        // the device fills slots ahead; the host empties slot k once it is
        // filled, and the slot is refilled once the host is done with it
        for (int k = 0; k < std::min(nch, kSlots) && e == hipSuccess; k++) e = issue(k);
        for (int k = 0; k < nch && e == hipSuccess; k++) {
            e = hipEventSynchronize(ev[k % kSlots]);
            if (e != hipSuccess) break;
            host(k);
            pool_wait(grp[k]);
            if (k + kSlots < nch) e = issue(k + kSlots);
        }
    }
    // the items read and write the caller's memory and the slots: all of
    // them have run before returning, also after an error
    for (int k = 0; k < submitted; k++) pool_wait(grp[k]);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

}  // namespace hz
