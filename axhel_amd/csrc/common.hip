// Host-side helpers: error text, device queries, host-pointer staging.
#include "common.hpp"

#include <cstdio>

namespace axhel_host {

static thread_local std::string g_last_error;

void set_last_error(const char *where, hipError_t err) {
    g_last_error = std::string(where) + ": " + hipGetErrorString(err);
}

void set_last_error_text(const std::string &text) { g_last_error = text; }

bool is_device_pointer(const void *p) {
    if (p == nullptr) return false;
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory: clear the error
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int cu_count() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cache[dev] = v;
    }
    return cache[dev];
}

namespace {

// Idle staging buffers by device, sizes rounded up to powers of two (64 KiB
// at least).  A per-call hipMalloc / hipFree costs more than the transform of
// a limb (hipFree waits for the device), so buffers are kept; at most
// kPoolCap bytes stay idle.  The pool is never destroyed: no HIP call runs
// during static destruction at exit.
struct StagePool {
    struct Buf {
        int device;
        size_t bytes;
        void *p;
    };
    std::mutex mu;
    std::vector<Buf> idle;
    size_t idle_bytes = 0;
};

constexpr size_t kPoolCap = size_t(256) << 20;

StagePool &stage_pool() {
    static StagePool *pool = new StagePool;
    return *pool;
}

size_t pooled_size(size_t bytes) {
    size_t b = size_t(64) << 10;
    while (b < bytes) b <<= 1;
    return b;
}

hipError_t pool_acquire(int device, size_t bytes, void **p) {
    StagePool &pool = stage_pool();
    {
        std::lock_guard<std::mutex> lk(pool.mu);
        for (size_t i = 0; i < pool.idle.size(); ++i) {
            if (pool.idle[i].device == device && pool.idle[i].bytes == bytes) {
                *p = pool.idle[i].p;
                pool.idle_bytes -= bytes;
                pool.idle[i] = pool.idle.back();
                pool.idle.pop_back();
                return hipSuccess;
            }
        }
    }
    return hipMalloc(p, bytes);
}

void pool_release(int device, size_t bytes, void *p) {
    StagePool &pool = stage_pool();
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lk(pool.mu);
        pool.idle.push_back({device, bytes, p});
        pool.idle_bytes += bytes;
        while (pool.idle_bytes > kPoolCap && !pool.idle.empty()) {  // oldest first
            drop.push_back(pool.idle.front().p);
            pool.idle_bytes -= pool.idle.front().bytes;
            pool.idle.erase(pool.idle.begin());
        }
    }
    for (void *d : drop) (void)hipFree(d);
}

}  // namespace

Staging::~Staging() {
    if (entries_.empty()) return;
    // An error path may leave copies or kernels queued on the buffers.  If
    // even this wait fails the stream may still own them: they are dropped,
    // not pooled (a device error ends the process's GPU work anyway).
    if (!synced_ && hipStreamSynchronize(stream_) != hipSuccess) return;
    for (auto &e : entries_) pool_release(e.device, e.bytes, e.dev);
}

hipError_t Staging::stage(const void *p, uint64_t n, bool copy_in, void **dev, Entry **entry) {
    for (auto &e : entries_) {
        if (e.host == p && e.n == n) {
            *dev = e.dev;
            *entry = &e;
            return hipSuccess;
        }
    }
    int device = 0;
    hipError_t err = hipGetDevice(&device);
    if (err != hipSuccess) return err;
    const size_t bytes = pooled_size(n * sizeof(uint64_t));
    void *d = nullptr;
    err = pool_acquire(device, bytes, &d);
    if (err != hipSuccess) return err;
    // registered before the copy, so every exit path returns it to the pool
    entries_.push_back(Entry{p, d, n, false, bytes, device});
    if (copy_in && n) {
        err = hipMemcpyAsync(d, p, n * sizeof(uint64_t), hipMemcpyHostToDevice, stream_);
        if (err != hipSuccess) return err;
    }
    *dev = d;
    *entry = &entries_.back();
    return hipSuccess;
}

hipError_t Staging::input(const uint64_t *p, uint64_t n, const uint64_t **dev) {
    if (p == nullptr || is_device_pointer(p)) {
        *dev = p;
        return hipSuccess;
    }
    void *d = nullptr;
    Entry *e = nullptr;
    hipError_t err = stage(p, n, true, &d, &e);
    *dev = static_cast<const uint64_t *>(d);
    return err;
}

hipError_t Staging::output(uint64_t *p, uint64_t n, uint64_t **dev) {
    if (p == nullptr || is_device_pointer(p)) {
        *dev = p;
        return hipSuccess;
    }
    void *d = nullptr;
    Entry *e = nullptr;
    // The result buffer is fully overwritten by every kernel, so a fresh
    // staging buffer needs no copy-in unless it aliases an operand (handled
    // by stage() returning the operand's buffer).
    hipError_t err = stage(p, n, false, &d, &e);
    if (err == hipSuccess) e->is_out = true;
    *dev = static_cast<uint64_t *>(d);
    return err;
}

hipError_t Staging::finish() {
    if (entries_.empty()) return hipSuccess;
    for (auto &e : entries_) {
        if (!e.is_out || e.n == 0) continue;
        hipError_t err = hipMemcpyAsync(const_cast<void *>(e.host), e.dev, e.n * sizeof(uint64_t),
                                        hipMemcpyDeviceToHost, stream_);
        if (err != hipSuccess) return err;
    }
    const hipError_t err = hipStreamSynchronize(stream_);
    synced_ = err == hipSuccess;
    return err;
}

}  // namespace axhel_host

// ------------------------------------------------------------------ C ABI

extern "C" {

const char *axhel_hip_status_string(axhel_hip_status s) {
    switch (s) {
        case AXHEL_HIP_OK: return "ok";
        case AXHEL_HIP_UNSUPPORTED_FACTOR: return "unsupported mod factor (no-op)";
        case AXHEL_HIP_INVALID_ARGUMENT: return "invalid argument";
        case AXHEL_HIP_NO_MEMORY: return "out of device memory";
        case AXHEL_HIP_HIP_ERROR: return "HIP runtime error";
        case AXHEL_HIP_NO_DEVICE: return "no GPU device";
        case AXHEL_HIP_REJECTED: return "request rejected";
        case AXHEL_HIP_TIMEOUT: return "timed out";
        case AXHEL_HIP_PENDING: return "still pending";
        case AXHEL_HIP_CANCELLED: return "cancelled";
    }
    return "unknown status";
}

const char *axhel_hip_last_error(void) { return axhel_host::g_last_error.c_str(); }

int axhel_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int axhel_hip_is_device_pointer(const void *p) { return axhel_host::is_device_pointer(p) ? 1 : 0; }

axhel_hip_status axhel_hip_stream_synchronize(void *stream) {
    AXHEL_HIP_TRY(hipStreamSynchronize(axhel_host::as_stream(stream)));
    return AXHEL_HIP_OK;
}

}  // extern "C"
