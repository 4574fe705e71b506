// Large host <-> device copies between pageable caller memory and the device
// through three 16 MB slots of pinned memory: asynchronous copies move whole
// slots on the stream while up to 8 host threads copy between the caller's
// memory and the other slots.  Pinned copies run near the PCIe rate and the
// host side is parallel (it also takes the first-touch page faults of a
// freshly allocated destination in parallel), where a plain copy from / to
// pageable memory goes through one runtime staging buffer.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
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
    unsigned hw = std::thread::hardware_concurrency();
    const int nt = (int)std::max(1u, std::min(hw ? hw : 1u, 8u));
    std::atomic<int> go{-1};  // host side of chunks <= go may start
    std::atomic<bool> abort{false};
    std::vector<std::atomic<int>> done(nch);
    for (auto& x : done) x.store(0);
    std::vector<std::thread> th;
    auto work = [&](int t) {
        for (int k = 0; k < nch; k++) {
            while (go.load(std::memory_order_acquire) < k) {
                if (abort.load()) return;
                std::this_thread::yield();
            }
            const size_t n = len(k), b = n / nt * t, en = t == nt - 1 ? n : n / nt * (t + 1);
            if (to_device)
                std::memcpy(sl(k) + b, s + (size_t)k * kChunk + b, en - b);
            else
                std::memcpy(d + (size_t)k * kChunk + b, sl(k) + b, en - b);
            done[k].fetch_add(1, std::memory_order_release);
        }
    };
    try {
        for (int t = 0; t < nt; t++) th.emplace_back(work, t);
    } catch (...) {  // no threads: nothing was copied yet, copy directly
        abort.store(true);
        for (auto& x : th) x.join();
        return plain();
    }
    auto wait_done = [&](int k) {
        while (done[k].load(std::memory_order_acquire) < nt) std::this_thread::yield();
    };
    if (to_device) {
        // host fills slot k (once its previous device copy is over), then the
        // device copies it
        for (int k = 0; k < nch && e == hipSuccess; k++) {
            if (k >= kSlots) e = hipEventSynchronize(ev[k % kSlots]);
            if (e != hipSuccess) break;
            go.store(k, std::memory_order_release);
            wait_done(k);
            e = issue(k);
        }
    } else {
        // the device fills slots ahead; the host empties slot k once it is
        // filled, and the slot is refilled once the host is done with it
        for (int k = 0; k < std::min(nch, kSlots) && e == hipSuccess; k++) e = issue(k);
        for (int k = 0; k < nch && e == hipSuccess; k++) {
            e = hipEventSynchronize(ev[k % kSlots]);
            if (e != hipSuccess) break;
            go.store(k, std::memory_order_release);
            if (k + kSlots < nch) {
                wait_done(k);
                e = issue(k + kSlots);
            }
        }
    }
    if (e != hipSuccess) abort.store(true);
    for (auto& x : th) x.join();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

}  // namespace hz
