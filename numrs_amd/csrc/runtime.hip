// Runtime part of the C ABI: device selection, errors, memory, streams, events, and the
// per-stream scratch that the reduction kernels use for their cross-workgroup finish.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

namespace nrk {

static thread_local char g_last_error[512] = "";
static thread_local int g_thread_device = -1;

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

nrk_status fail(nrk_status code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

nrk_status hip_fail(hipError_t e, const char *what, const char *file, int line) {
    (void)hipGetLastError();  // clear the sticky non-fatal error state
    set_error("%s failed: %s (%d) at %s:%d", what, hipGetErrorString(e), (int)e, file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return NRK_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return NRK_ERR_NO_DEVICE;
    return NRK_ERR_HIP;
}

struct DeviceInfo {
    bool checked = false;
    bool ok = false;
    int cus = 0;
    void *scratch = nullptr;  // this device's copy of `scratch` (below)
};

// Fixed device scratch for nrk_dev_map_reduce_f64_noalloc, whose callers cannot allocate. Each
// device holds its own copy until the process ends. In host code `scratch` is only a handle: its
// device address comes from hipGetSymbolAddress with that device current (nrk_init), and only that
// address is ever passed to kernels.
static __device__ __align__(256) unsigned char scratch[NRK_NOALLOC_SCRATCH_BYTES];
static std::mutex g_dev_mutex;
static std::map<int, DeviceInfo> g_devices;

static nrk_status check_device(int device) {
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    DeviceInfo &info = g_devices[device];
    if (info.checked) {
        return info.ok ? NRK_OK : fail(NRK_ERR_NO_DEVICE, "device %d is not a gfx950 device", device);
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(NRK_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(NRK_ERR_NO_DEVICE, "device %d out of range [0,%d)", device, count);
    hipDeviceProp_t prop;
    NRK_HIP(hipGetDeviceProperties(&prop, device));
    info.checked = true;
    info.ok = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
    info.cus = prop.multiProcessorCount;
    if (!info.ok) return fail(NRK_ERR_NO_DEVICE, "device %d is %s, this library is built for gfx950", device, prop.gcnArchName);
    return NRK_OK;
}

// Looks up this device's address of `scratch` once (the device must be current).
static nrk_status init_static_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    DeviceInfo &info = g_devices[device];
    if (info.scratch) return NRK_OK;
    void *p = nullptr;
    NRK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(NRK_ERR_HIP, "hipGetSymbolAddress gave no address for the fixed scratch of device %d", device);
    info.scratch = p;
    return NRK_OK;
}

nrk_status static_scratch(ExtScratch *out) {
    NRK_TRY(ensure_device());
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    auto it = g_devices.find(g_thread_device);
    if (it == g_devices.end() || !it->second.scratch) return fail(NRK_ERR_HIP, "fixed scratch of device %d is not set up", g_thread_device);
    out->base = it->second.scratch;
    out->bytes = NRK_NOALLOC_SCRATCH_BYTES;
    return NRK_OK;
}

nrk_status ensure_device() {
    if (g_thread_device >= 0) return NRK_OK;
    return nrk_init(0);
}

int cu_count() {
    int dev = g_thread_device < 0 ? 0 : g_thread_device;
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    auto it = g_devices.find(dev);
    return (it != g_devices.end() && it->second.cus > 0) ? it->second.cus : 256;
}

// ---- the library's device allocations ----
// Every buffer the library allocates for itself (per-stream scratch and workspace, host-call
// staging, nrk_device_malloc) goes through lib_alloc: an out-of-memory result becomes
// NRK_ERR_OUT_OF_MEMORY with the size and purpose, and nrk_debug_fail_allocations can make
// the n-th next one fail (a test hook for the error paths; one failure, then it disarms).
static std::atomic<long long> g_fail_after{-1};

static nrk_status lib_alloc(void **p, size_t bytes, const char *what) {
    *p = nullptr;
    long long k = g_fail_after.load();
    while (k >= 0 && !g_fail_after.compare_exchange_weak(k, k - 1)) {
    }
    if (k == 0) return fail(NRK_ERR_OUT_OF_MEMORY, "device allocation of %zu bytes for %s failed (injected)", bytes, what);
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(NRK_ERR_OUT_OF_MEMORY, "device allocation of %zu bytes for %s failed: out of memory", bytes, what);
    }
    NRK_HIP(e);
    return NRK_OK;
}

// ---- per-(device, stream) scratch ----
struct StreamScratch {
    unsigned int *counters = nullptr;
    size_t n_counters = 0;
    void *partials = nullptr;
    size_t partial_bytes = 0;
};
static std::mutex g_scratch_mutex;
static std::map<std::pair<int, hipStream_t>, StreamScratch> g_scratch;

nrk_status get_scratch(hipStream_t stream, size_t n_counters, size_t partial_bytes, Scratch *out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    StreamScratch &s = g_scratch[{g_thread_device, stream}];
    if (n_counters > s.n_counters) {
        size_t want = align_up(n_counters < 4096 ? 4096 : n_counters, 4096);
        if (s.counters) {
            NRK_HIP(hipStreamSynchronize(stream));
            NRK_HIP(hipFree(s.counters));
            s.counters = nullptr;
            s.n_counters = 0;
        }
        NRK_TRY(lib_alloc((void **)&s.counters, want * sizeof(unsigned int), "reduction counters"));
        // Establishes the zero invariant once; kernels keep it.
        NRK_HIP(hipMemset(s.counters, 0, want * sizeof(unsigned int)));
        s.n_counters = want;
    }
    if (partial_bytes > s.partial_bytes) {
        size_t want = align_up(partial_bytes < (1u << 20) ? (1u << 20) : partial_bytes, 1u << 20);
        if (s.partials) {
            NRK_HIP(hipStreamSynchronize(stream));
            NRK_HIP(hipFree(s.partials));
            s.partials = nullptr;
            s.partial_bytes = 0;
        }
        NRK_TRY(lib_alloc(&s.partials, want, "reduction scratch"));
        s.partial_bytes = want;
    }
    out->counters = s.counters;
    out->n_counters = s.n_counters;
    out->partials = s.partials;
    out->partial_bytes = s.partial_bytes;
    return NRK_OK;
}

// ---- per-(device, stream) workspace ----
struct StreamWorkspace {
    void *ptr = nullptr;
    size_t bytes = 0;
};
static std::map<std::pair<int, hipStream_t>, StreamWorkspace> g_workspace;

nrk_status get_workspace(hipStream_t stream, size_t bytes, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    StreamWorkspace &w = g_workspace[{g_thread_device, stream}];
    if (bytes > w.bytes) {
        if (w.ptr) {
            NRK_HIP(hipStreamSynchronize(stream));
            NRK_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        const size_t want = align_up(bytes, 1u << 21);
        NRK_TRY(lib_alloc(&w.ptr, want, "stream workspace"));
        w.bytes = want;
    }
    *out = w.ptr;
    return NRK_OK;
}

// ---- host <-> device copies of the host-pointer entry points ----
// A copy from or to pageable memory goes through a ring of pinned buffers: the host side of each
// chunk (pageable <-> pinned) is copied by a small pool of threads while the DMA of the previous
// chunk runs. Measured on the MI355X host: pageable hipMemcpy 27-28 GB/s to the device and 17 GB/s
// back, pinned DMA 57 GB/s either way, pageable -> pinned memcpy 29 GB/s on one thread and 87 GB/s
// on four.
constexpr size_t STAGE_MIN = (size_t)1 << 20;     // below this, one plain hipMemcpyAsync
constexpr size_t STAGE_CHUNK = (size_t)8 << 20;   // bytes per ring buffer
constexpr int STAGE_NBUF = 4;                      // ring buffers per host thread and device
constexpr int COPY_PARTS = 4;                      // host copy: the caller and three pool threads

class CopyPool {
  public:
    static CopyPool &get() {
        static CopyPool pool;
        return pool;
    }
    // memcpy(d, s, n), split over the pool; returns when every part is done
    void copy(void *d, const void *s, size_t n) {
        std::lock_guard<std::mutex> one(run_mu_);  // one parallel copy at a time
        if (threads_.empty())
            for (int i = 0; i < COPY_PARTS - 1; ++i) threads_.emplace_back([this, i] { worker(i); });
        {
            std::lock_guard<std::mutex> l(mu_);
            dst_ = (char *)d;
            src_ = (const char *)s;
            n_ = n;
            pending_ = COPY_PARTS - 1;
            ++gen_;
        }
        work_cv_.notify_all();
        part(COPY_PARTS - 1);
        std::unique_lock<std::mutex> l(mu_);
        done_cv_.wait(l, [this] { return pending_ == 0; });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
        }
        work_cv_.notify_all();
        for (auto &t : threads_) t.join();
    }

  private:
    void part(int i) {  // part i of COPY_PARTS, cut on 64-byte boundaries
        const size_t lo = (n_ * i / COPY_PARTS) & ~(size_t)63, hi = i + 1 == COPY_PARTS ? n_ : (n_ * (i + 1) / COPY_PARTS) & ~(size_t)63;
        if (hi > lo) std::memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void worker(int i) {
        unsigned seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> l(mu_);
            work_cv_.wait(l, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            l.unlock();
            part(i);
            l.lock();
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }
    std::mutex run_mu_, mu_;
    std::condition_variable work_cv_, done_cv_;
    std::vector<std::thread> threads_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t n_ = 0;
    unsigned gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

static void host_copy(void *d, const void *s, size_t n) {
    if (n < STAGE_MIN) std::memcpy(d, s, n);
    else CopyPool::get().copy(d, s, n);
}

struct StageRing {
    bool tried = false;  // set up once
    bool ok = false;     // pinned buffers and events exist; else copies take plain hipMemcpy
    char *buf[STAGE_NBUF] = {};
    hipEvent_t ev[STAGE_NBUF] = {};
    bool used[STAGE_NBUF] = {};
};
static thread_local std::map<int, StageRing> g_rings;

// This thread's ring on its current device, created on first use (no pinned memory: ok = false).
static StageRing &stage_ring() {
    StageRing &r = g_rings[g_thread_device];
    if (r.tried) return r;
    r.tried = true;
    bool good = true;
    for (int b = 0; b < STAGE_NBUF && good; ++b) {
        good = hipHostMalloc((void **)&r.buf[b], STAGE_CHUNK, hipHostMallocDefault) == hipSuccess &&
               hipEventCreateWithFlags(&r.ev[b], hipEventDisableTiming) == hipSuccess;
    }
    if (!good) {
        (void)hipGetLastError();
        for (int b = 0; b < STAGE_NBUF; ++b) {
            if (r.buf[b]) (void)hipHostFree(r.buf[b]);
            if (r.ev[b]) (void)hipEventDestroy(r.ev[b]);
            r.buf[b] = nullptr;
            r.ev[b] = nullptr;
        }
        return r;  // plain copies from now on
    }
    r.ok = true;
    return r;
}

nrk_status copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return NRK_OK;
    StageRing *r = bytes < STAGE_MIN ? nullptr : &stage_ring();
    if (!r || !r->ok) {
        NRK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return NRK_OK;
    }
    for (size_t off = 0, c = 0; off < bytes; ++c) {
        const size_t len = std::min(STAGE_CHUNK, bytes - off);
        const int b = (int)(c % STAGE_NBUF);
        if (r->used[b]) NRK_HIP(hipEventSynchronize(r->ev[b]));  // its previous DMA is done
        host_copy(r->buf[b], (const char *)src + off, len);
        NRK_HIP(hipMemcpyAsync((char *)dst + off, r->buf[b], len, hipMemcpyHostToDevice, s));
        NRK_HIP(hipEventRecord(r->ev[b], s));
        r->used[b] = true;
        off += len;
    }
    return NRK_OK;
}

nrk_status copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return NRK_OK;
    StageRing *r = bytes < STAGE_MIN ? nullptr : &stage_ring();
    if (!r || !r->ok) {
        NRK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        NRK_HIP(hipStreamSynchronize(s));
        return NRK_OK;
    }
    const size_t nchunk = ceil_div(bytes, STAGE_CHUNK);
    auto issue = [&](size_t c) -> nrk_status {  // DMA of chunk c into its ring buffer
        const int b = (int)(c % STAGE_NBUF);
        if (r->used[b]) NRK_HIP(hipEventSynchronize(r->ev[b]));
        const size_t off = c * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - off);
        NRK_HIP(hipMemcpyAsync(r->buf[b], (const char *)src + off, len, hipMemcpyDeviceToHost, s));
        NRK_HIP(hipEventRecord(r->ev[b], s));
        r->used[b] = true;
        return NRK_OK;
    };
    for (size_t c = 0; c < std::min(nchunk, (size_t)STAGE_NBUF); ++c) NRK_TRY(issue(c));
    for (size_t c = 0; c < nchunk; ++c) {
        const int b = (int)(c % STAGE_NBUF);
        NRK_HIP(hipEventSynchronize(r->ev[b]));
        const size_t off = c * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - off);
        host_copy((char *)dst + off, r->buf[b], len);
        if (c + STAGE_NBUF < nchunk) NRK_TRY(issue(c + STAGE_NBUF));
    }
    return NRK_OK;
}

// ---- host staging buffers (per thread, per slot) ----
static thread_local Staging g_staging[4];

nrk_status staging(int slot, size_t bytes, void **ptr) {
    Staging &s = g_staging[slot];
    if (bytes > s.bytes) {
        if (s.ptr) {
            NRK_HIP(hipDeviceSynchronize());
            NRK_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        size_t want = align_up(bytes < 4096 ? 4096 : bytes, 4096);
        NRK_TRY(lib_alloc(&s.ptr, want, "host-call staging"));
        s.bytes = want;
    }
    *ptr = s.ptr;
    return NRK_OK;
}

nrk_status make_chain(const nrk_scalar_op *chain, int32_t chain_len, Chain *out, bool allow_math) {
    if (chain_len < 0 || chain_len > NRK_MAX_CHAIN) return fail(NRK_ERR_INVALID_ARGUMENT, "chain_len %d outside [0,%d]", chain_len, NRK_MAX_CHAIN);
    if (chain_len > 0 && chain == nullptr) return fail(NRK_ERR_INVALID_ARGUMENT, "chain is NULL");
    out->len = chain_len;
    for (int i = 0; i < NRK_MAX_CHAIN; ++i) {
        out->op[i] = -1;
        out->s[i] = 0.0;
    }
    for (int i = 0; i < chain_len; ++i) {
        if (chain[i].op < 0 || chain[i].op >= NRK_SOP_COUNT) return fail(NRK_ERR_INVALID_ARGUMENT, "chain[%d].op = %d is not an NRK_SOP_* code", i, chain[i].op);
        if (!allow_math && chain[i].op >= NRK_SOP_EXP)
            return fail(NRK_ERR_UNSUPPORTED, "chain[%d].op = %d (an elementary function) runs in the elementwise map only", i, chain[i].op);
        out->op[i] = chain[i].op;
        out->s[i] = chain[i].scalar;
    }
    return NRK_OK;
}

}  // namespace nrk

using namespace nrk;

extern "C" {

nrk_status nrk_init(int device) {
    NRK_TRY(check_device(device));
    NRK_HIP(hipSetDevice(device));
    NRK_TRY(init_static_scratch(device));
    g_thread_device = device;
    return NRK_OK;
}

const char *nrk_status_string(nrk_status status) {
    switch (status) {
        case NRK_OK: return "ok";
        case NRK_ERR_INVALID_ARGUMENT: return "invalid argument";
        case NRK_ERR_NO_DEVICE: return "no usable gfx950 device";
        case NRK_ERR_OUT_OF_MEMORY: return "out of device memory";
        case NRK_ERR_HIP: return "HIP runtime error";
        case NRK_ERR_UNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

const char *nrk_last_error(void) { return g_last_error; }
const char *nrk_version(void) { return "0.1.0"; }

nrk_status nrk_device_cu_count(int32_t *count) {
    if (!count) return fail(NRK_ERR_INVALID_ARGUMENT, "count is NULL");
    NRK_TRY(ensure_device());
    *count = cu_count();
    return NRK_OK;
}

nrk_status nrk_device_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(NRK_ERR_INVALID_ARGUMENT, "ptr is NULL");
    NRK_TRY(ensure_device());
    *ptr = nullptr;
    if (bytes == 0) bytes = 16;
    NRK_TRY(lib_alloc(ptr, bytes, "nrk_device_malloc"));
    return NRK_OK;
}

nrk_status nrk_debug_fail_allocations(int32_t after) {
    g_fail_after.store(after < 0 ? -1 : (long long)after);
    return NRK_OK;
}

nrk_status nrk_device_free(void *ptr) {
    if (!ptr) return NRK_OK;
    NRK_HIP(hipFree(ptr));
    return NRK_OK;
}

nrk_status nrk_memcpy_htod(void *dst, const void *src, size_t bytes, nrk_stream stream) {
    if (bytes == 0) return NRK_OK;
    NRK_TRY(ensure_device());
    NRK_TRY(copy_h2d(dst, src, bytes, (hipStream_t)stream));
    NRK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return NRK_OK;
}

nrk_status nrk_memcpy_dtoh(void *dst, const void *src, size_t bytes, nrk_stream stream) {
    if (bytes == 0) return NRK_OK;
    NRK_TRY(ensure_device());
    return copy_d2h(dst, src, bytes, (hipStream_t)stream);
}

nrk_status nrk_memcpy_dtod(void *dst, const void *src, size_t bytes, nrk_stream stream) {
    if (bytes == 0) return NRK_OK;
    NRK_TRY(ensure_device());
    NRK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return NRK_OK;
}

nrk_status nrk_memset_zero(void *dst, size_t bytes, nrk_stream stream) {
    if (bytes == 0) return NRK_OK;
    NRK_TRY(ensure_device());
    NRK_HIP(hipMemsetAsync(dst, 0, bytes, (hipStream_t)stream));
    return NRK_OK;
}

nrk_status nrk_stream_create(nrk_stream *stream) {
    if (!stream) return fail(NRK_ERR_INVALID_ARGUMENT, "stream is NULL");
    NRK_TRY(ensure_device());
    hipStream_t s;
    NRK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (nrk_stream)s;
    return NRK_OK;
}

nrk_status nrk_stream_destroy(nrk_stream stream) {
    if (!stream) return NRK_OK;
    NRK_HIP(hipStreamSynchronize((hipStream_t)stream));
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        auto it = g_scratch.find({g_thread_device, (hipStream_t)stream});
        if (it != g_scratch.end()) {
            if (it->second.counters) (void)hipFree(it->second.counters);
            if (it->second.partials) (void)hipFree(it->second.partials);
            g_scratch.erase(it);
        }
        auto wt = g_workspace.find({g_thread_device, (hipStream_t)stream});
        if (wt != g_workspace.end()) {
            if (wt->second.ptr) (void)hipFree(wt->second.ptr);
            g_workspace.erase(wt);
        }
    }
    NRK_HIP(hipStreamDestroy((hipStream_t)stream));
    return NRK_OK;
}

nrk_status nrk_stream_synchronize(nrk_stream stream) {
    NRK_TRY(ensure_device());
    NRK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return NRK_OK;
}

nrk_status nrk_device_synchronize(void) {
    NRK_TRY(ensure_device());
    NRK_HIP(hipDeviceSynchronize());
    return NRK_OK;
}

nrk_status nrk_event_create(void **event) {
    if (!event) return fail(NRK_ERR_INVALID_ARGUMENT, "event is NULL");
    NRK_TRY(ensure_device());
    hipEvent_t e;
    NRK_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return NRK_OK;
}

nrk_status nrk_event_destroy(void *event) {
    if (!event) return NRK_OK;
    NRK_HIP(hipEventDestroy((hipEvent_t)event));
    return NRK_OK;
}

nrk_status nrk_event_record(void *event, nrk_stream stream) {
    NRK_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return NRK_OK;
}

nrk_status nrk_stream_wait_event(nrk_stream stream, void *event) {
    NRK_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return NRK_OK;
}

nrk_status nrk_event_elapsed_ms(float *ms, void *start, void *stop) {
    if (!ms) return fail(NRK_ERR_INVALID_ARGUMENT, "ms is NULL");
    NRK_HIP(hipEventSynchronize((hipEvent_t)stop));
    NRK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return NRK_OK;
}

}  // extern "C"
