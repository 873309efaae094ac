// xfer.hip — host<->device copies of caller-owned (pageable) host memory.
//
// Large transfers are staged through two pinned buffers of XFER_CHUNK bytes that the library
// allocates once (hipHostMalloc) and reuses; copies of up to XFER_DIRECT bytes go directly.  The
// runtime is never asked to pin the caller's memory.  Both functions return when the data has
// arrived.

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "rbtp_internal.h"

namespace rbtp {

namespace {

constexpr size_t XFER_DIRECT = (size_t)1 << 20;
constexpr size_t XFER_CHUNK = (size_t)16 << 20;

struct Staging {
    std::mutex mu;
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int device = -1;
};

Staging &staging() {
    static Staging *s = new Staging();  // kept until the process ends
    return *s;
}

// Pinned buffers and events for the current device (reallocated if the device changed).
hipError_t staging_ready(Staging &s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (s.buf[0] && s.device == dev) return hipSuccess;
    for (int b = 0; b < 2; b++) {
        if (s.buf[b]) (void)hipHostFree(s.buf[b]);
        if (s.ev[b]) (void)hipEventDestroy(s.ev[b]);
        s.buf[b] = nullptr;
        s.ev[b] = nullptr;
    }
    for (int b = 0; b < 2 && e == hipSuccess; b++) {
        e = hipHostMalloc((void **)&s.buf[b], XFER_CHUNK, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev[b], hipEventDisableTiming);
    }
    if (e == hipSuccess) s.device = dev;
    return e;
}

}  // namespace

hipError_t copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    if (bytes <= XFER_DIRECT) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
        return e == hipSuccess ? hipStreamSynchronize(st) : e;
    }
    Staging &s = staging();
    std::lock_guard<std::mutex> lock(s.mu);
    hipError_t e = staging_ready(s);
    bool used[2] = {false, false};
    for (size_t off = 0, k = 0; off < bytes && e == hipSuccess; off += XFER_CHUNK, k++) {
        const int b = (int)(k & 1);
        const size_t len = bytes - off < XFER_CHUNK ? bytes - off : XFER_CHUNK;
        if (used[b]) e = hipEventSynchronize(s.ev[b]);  // the buffer's previous copy is done
        if (e != hipSuccess) break;
        std::memcpy(s.buf[b], (const char *)src + off, len);
        e = hipMemcpyAsync((char *)dst + off, s.buf[b], len, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(s.ev[b], st);
        used[b] = true;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

hipError_t copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    if (bytes <= XFER_DIRECT) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st);
        return e == hipSuccess ? hipStreamSynchronize(st) : e;
    }
    Staging &s = staging();
    std::lock_guard<std::mutex> lock(s.mu);
    hipError_t e = staging_ready(s);
    size_t pend_off = 0, pend_len = 0;
    int pend_b = -1;
    for (size_t off = 0, k = 0; off < bytes && e == hipSuccess; off += XFER_CHUNK, k++) {
        const int b = (int)(k & 1);
        const size_t len = bytes - off < XFER_CHUNK ? bytes - off : XFER_CHUNK;
        e = hipMemcpyAsync(s.buf[b], (const char *)src + off, len, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(s.ev[b], st);
        if (e == hipSuccess && pend_b >= 0) {  // drain the previous chunk while this one flies
            e = hipEventSynchronize(s.ev[pend_b]);
            if (e == hipSuccess) std::memcpy((char *)dst + pend_off, s.buf[pend_b], pend_len);
        }
        pend_b = b;
        pend_off = off;
        pend_len = len;
    }
    if (e == hipSuccess && pend_b >= 0) {
        e = hipEventSynchronize(s.ev[pend_b]);
        if (e == hipSuccess) std::memcpy((char *)dst + pend_off, s.buf[pend_b], pend_len);
    }
    return e;
}

}  // namespace rbtp
