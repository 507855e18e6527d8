// Host-side helpers shared by the HIP translation units: status plumbing,
// per-thread last-error text, device queries and host-pointer staging.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/axhel_hip.h"

namespace axhel_host {

// Records the HIP error text for axhel_hip_last_error() on this thread.
void set_last_error(const char *where, hipError_t err);
void set_last_error_text(const std::string &text);

#define AXHEL_HIP_TRY(expr)                                        \
    do {                                                           \
        hipError_t _e = (expr);                                    \
        if (_e != hipSuccess) {                                    \
            ::axhel_host::set_last_error(#expr, _e);               \
            return _e == hipErrorOutOfMemory ? AXHEL_HIP_NO_MEMORY \
                                             : AXHEL_HIP_HIP_ERROR; \
        }                                                          \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// True if p is device-accessible memory (device or managed allocation).
bool is_device_pointer(const void *p);

// Number of compute units of the current device (cached per device).
int cu_count();

// 16-byte alignment test for vectorised access paths.
inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Staging of host operands for the synchronous host-pointer path of the
// entry points.  Device pointers pass through untouched; each distinct host
// pointer gets one device buffer (exact aliasing between an operand and the
// result therefore shares a buffer, as in-place calls need).  Buffers come
// from a process-wide pool and go back to it once the call's stream has
// synchronised, so a steady stream of host-pointer calls allocates nothing.
class Staging {
public:
    explicit Staging(hipStream_t s) : stream_(s) {}
    ~Staging();
    Staging(const Staging &) = delete;
    Staging &operator=(const Staging &) = delete;

    hipError_t input(const uint64_t *p, uint64_t n, const uint64_t **dev);
    hipError_t output(uint64_t *p, uint64_t n, uint64_t **dev);
    // Copies staged outputs back to their host pointers, then synchronises
    // the stream when anything was staged.
    hipError_t finish();
    bool any() const { return !entries_.empty(); }

private:
    struct Entry {
        const void *host;
        void *dev;
        uint64_t n;
        bool is_out;
        size_t bytes;  // pooled size
        int device;
    };
    hipError_t stage(const void *p, uint64_t n, bool copy_in, void **dev, Entry **entry);
    std::vector<Entry> entries_;
    hipStream_t stream_;
    bool synced_ = false;  // finish() synchronised the stream: buffers are idle
};

}  // namespace axhel_host
