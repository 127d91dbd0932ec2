// Device context: one HIP stream, memory helpers and per-kind launch timers.
#include "common.h"

#include <string.h>

#include <mutex>
#include <time.h>
#include <unordered_map>
#include <vector>

namespace irwls {

static bool g_init = false;
static hipStream_t g_stream = nullptr;
// A host thread of a batched run works on a stream of its own (and its own
// scratch, staging ring and solver buffers, all thread-local below).
static thread_local hipStream_t t_stream = nullptr;

hipStream_t stream() {
    if (t_stream) return t_stream;
    if (!g_init) irwls_dev_init();
    return g_stream;
}

bool dev_initialised() { return g_init; }

// ---- timers --------------------------------------------------------------
// Each timed launch records an event pair from a pool without synchronising;
// pairs are resolved (elapsed time read) when the pool fills or the host asks
// for the totals, so timing does not serialise the stream.
static bool g_timing = false;
static unsigned g_timing_mask = ~0u;   // kinds that are timed while timing is on
// Every g_stride-th launch scope of a timed kind carries events (1: all).
// The others run unobserved, and their work is not counted either, so a
// kind's time, launches, flops and bytes describe the same sampled launches.
static int g_stride = 1;
static long g_seen[IRWLS_T_COUNT];
static long g_launch[IRWLS_T_COUNT];
static double g_ms[IRWLS_T_COUNT], g_flops[IRWLS_T_COUNT], g_bytes[IRWLS_T_COUNT];
static const int kPool = 4096;
static hipEvent_t g_evs[kPool][2];
static int g_evkind[kPool];
static bool g_evdone[kPool];
static int g_evn = 0;
static bool g_pool_ready = false;
// Timed launches of several host threads (batched solves of different round
// slots) may overlap: each reserves its event pair under this lock.
static std::mutex g_tmu;

// Reads the elapsed times of the recorded pairs; with g_tmu held, and only
// when every reserved pair has been recorded.  One device-wide wait instead
// of a wait per event: the streams the pairs were recorded on may have been
// destroyed since (worker threads end theirs).
static bool resolve_pool_locked() {
    if (g_evn == 0) return true;
    for (int i = 0; i < g_evn; i++)
        if (!g_evdone[i]) return false;
    IRWLS_HIP(hipDeviceSynchronize());
    for (int i = 0; i < g_evn; i++) {
        if (g_evkind[i] < 0) continue;
        float ms = 0.f;
        IRWLS_HIP(hipEventElapsedTime(&ms, g_evs[i][0], g_evs[i][1]));
        g_ms[g_evkind[i]] += ms;
    }
    g_evn = 0;
    return true;
}

static void resolve_pool() {
    std::lock_guard<std::mutex> lk(g_tmu);
    resolve_pool_locked();
}

hipEvent_t timer_event(int idx, int which) { return g_evs[idx][which]; }

TimerScope::TimerScope(int k, double f, double b, bool any_thread, bool mk)
    : kind(k), flops(f), bytes(b),
      on(k >= 0 && g_timing && (any_thread || !t_stream) && ((g_timing_mask >> k) & 1u)), markers(mk),
      idx(-1), nk(0) {
    if (!on) return;
    std::lock_guard<std::mutex> lk(g_tmu);
    if (g_stride > 1 && (g_seen[kind]++ % g_stride) != 0) {
        on = false;
        return;
    }
    if (g_evn == kPool && !resolve_pool_locked()) {
        on = false;   // pool full of pairs still in flight: this launch goes untimed
        return;
    }
    idx = g_evn++;
    g_evdone[idx] = false;
    if (markers) IRWLS_HIP(hipEventRecord(g_evs[idx][0], stream()));
}

TimerScope::~TimerScope() {
    if (!on) return;
    std::lock_guard<std::mutex> lk(g_tmu);
    const bool counted = markers || nk > 0;
    if (markers) {
        IRWLS_HIP(hipEventRecord(g_evs[idx][1], stream()));
    } else if (nk == 0) {
        // no kernel ran in the scope: an empty pair, not counted
        IRWLS_HIP(hipEventRecord(g_evs[idx][0], stream()));
        IRWLS_HIP(hipEventRecord(g_evs[idx][1], stream()));
    }
    g_evkind[idx] = counted ? kind : -1;
    g_evdone[idx] = true;
    if (!counted) return;
    g_launch[kind] += 1;
    g_flops[kind] += flops;
    g_bytes[kind] += bytes;
}

// Graph capture excludes every use of the timer events: HIP refuses an
// event operation while the stream the event was last recorded on captures.
void capture_lock() { g_tmu.lock(); }
void capture_unlock() { g_tmu.unlock(); }

int device_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, n = 0;
        IRWLS_HIP(hipGetDevice(&dev));
        IRWLS_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        cus = n > 0 ? n : 1;
    }
    return cus;
}

// ---- device memory cache -------------------------------------------------
// hipFree waits for the whole device, so with many worker threads (batched
// training) every free stalls its thread on every other thread's work.
// Freed blocks go back to per-size free lists instead.  The caller has
// synchronized the stream that used the block, so any thread may take it
// again at once.  Sizes are rounded up to 2^k x {1.25, 1.5, 1.75, 2} (at least
// 4 KiB).  The lists go back to the driver only when an allocation fails and
// when the library is trimmed or shut down.  Blocks above kCacheMax are
// returned at once (a buffer that grows in steps would otherwise keep every
// step).
static std::mutex g_cmu;
static std::unordered_map<void *, size_t> g_csize;            // every cached-allocator block: its size
static std::unordered_map<size_t, std::vector<void *>> g_cfree;
static const size_t kCacheMax = (size_t)1 << 30;

static size_t size_class(size_t b) {
    if (b <= 4096) return 4096;
    size_t p = 4096;
    while (p * 2 < b) p *= 2;   // p < b <= 2p
    for (size_t q = 5; q < 8; q++)
        if (p / 4 * q >= b) return p / 4 * q;
    return 2 * p;
}

static void cache_release_locked() {
    for (auto &kv : g_cfree)
        for (void *q : kv.second) {
            g_csize.erase(q);
            IRWLS_HIP(hipFree(q));
        }
    g_cfree.clear();
}

static void *cached_malloc(size_t bytes, size_t *got) {
    const size_t c = size_class(bytes);
    *got = c;
    {
        std::lock_guard<std::mutex> g(g_cmu);
        auto it = g_cfree.find(c);
        if (it != g_cfree.end() && !it->second.empty()) {
            void *q = it->second.back();
            it->second.pop_back();
            return q;
        }
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, c);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> g(g_cmu);
        cache_release_locked();
        e = hipMalloc(&q, c);
    }
    IRWLS_HIP(e);
    std::lock_guard<std::mutex> g(g_cmu);
    g_csize[q] = c;
    return q;
}

static void cached_free(void *q) {
    if (!q) return;
    {
        std::lock_guard<std::mutex> g(g_cmu);
        auto it = g_csize.find(q);
        if (it == g_csize.end()) {
            fprintf(stderr, "libirwls: freeing device memory the library did not allocate\n");
            exit(2);
        }
        if (it->second <= kCacheMax) {
            g_cfree[it->second].push_back(q);
            return;
        }
        g_csize.erase(it);
    }
    IRWLS_HIP(hipFree(q));
}

static void cache_release() {
    std::lock_guard<std::mutex> g(g_cmu);
    cache_release_locked();
}

// ---- scratch slots -------------------------------------------------------
static const int kSlots = 16;
static thread_local void *g_scr[kSlots];
static thread_local size_t g_scr_sz[kSlots];

void *scratch(int slot, size_t bytes) {
    if (slot < 0 || slot >= kSlots) {
        fprintf(stderr, "libirwls: bad scratch slot %d\n", slot);
        exit(2);
    }
    if (g_scr_sz[slot] < bytes) {
        if (g_scr[slot]) {
            IRWLS_HIP(hipStreamSynchronize(stream()));
            cached_free(g_scr[slot]);
        }
        g_scr[slot] = cached_malloc(bytes + (bytes >> 2) + 256, &g_scr_sz[slot]);
    }
    return g_scr[slot];
}

static void scratch_release() {
    for (int s = 0; s < kSlots; s++) {
        cached_free(g_scr[s]);
        g_scr[s] = nullptr;
        g_scr_sz[s] = 0;
    }
}

// ---- pinned host buffers and worker streams ------------------------------
// Worker threads of a batched run take their pinned staging buffers and
// their stream from pools and give them back when they finish: creating a
// stream and freeing pinned memory both cost milliseconds, and hipHostFree
// also waits for the device.
static std::mutex g_pmu;
static std::unordered_map<size_t, std::vector<void *>> g_pinned;
static std::vector<hipStream_t> g_streams;

static void *pinned_get(size_t bytes) {
    {
        std::lock_guard<std::mutex> g(g_pmu);
        auto it = g_pinned.find(bytes);
        if (it != g_pinned.end() && !it->second.empty()) {
            void *q = it->second.back();
            it->second.pop_back();
            return q;
        }
    }
    void *q = nullptr;
    IRWLS_HIP(hipHostMalloc(&q, bytes, hipHostMallocDefault));
    return q;
}

static void pinned_put(void *q, size_t bytes) {
    if (!q) return;
    std::lock_guard<std::mutex> g(g_pmu);
    auto &v = g_pinned[bytes];
    for (void *o : v)
        if (o == q) {
            fprintf(stderr, "libirwls: pinned buffer %p returned to its pool twice\n", q);
            exit(2);
        }
    v.push_back(q);
}

}  // namespace irwls

using namespace irwls;

// Plain device-to-device copy, 16 bytes per thread and step: the kernel of
// known traffic (reads = writes = the copied bytes) against which the
// hardware's fetch and write counters are calibrated.
__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void fill_zero_kernel(unsigned char *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    // 8-byte body when aligned, bytes otherwise
    if (((uintptr_t)p & 7) == 0) {
        size_t n8 = n / 8;
        uint64_t *q = (uint64_t *)p;
        for (size_t k = i; k < n8; k += stride) q[k] = 0;
        for (size_t k = n8 * 8 + i; k < n; k += stride) p[k] = 0;
    } else {
        for (size_t k = i; k < n; k += stride) p[k] = 0;
    }
}

extern "C" {

// The HIP runtime may draw from libc's random() state while it starts up.
// The training entry points reproduce the reference's rand() sequence from
// the caller's srand(), which the reference's drivers call before
// initMemory(), so start-up runs on a private state and the caller's state,
// position included, is put back afterwards.
static char g_rand_private[256];

int irwls_dev_init(void) {
    if (g_init) return 0;
    char *caller_state = initstate(1u, g_rand_private, sizeof(g_rand_private));
    int ndev = 0;
    IRWLS_HIP(hipGetDeviceCount(&ndev));
    if (ndev < 1) {
        fprintf(stderr, "libirwls: no GPU device visible\n");
        exit(2);
    }
    IRWLS_HIP(hipSetDevice(0));
    IRWLS_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    if (!g_pool_ready) {
        for (int i = 0; i < kPool; i++) {
            IRWLS_HIP(hipEventCreate(&g_evs[i][0]));
            IRWLS_HIP(hipEventCreate(&g_evs[i][1]));
        }
        g_pool_ready = true;
    }
    g_init = true;
    noalloc_setup();
    // first allocation and first launch (code-object load) inside the guard too
    void *probe = nullptr;
    IRWLS_HIP(hipMalloc(&probe, 256));
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(256), 0, g_stream, (unsigned char *)probe, (size_t)256);
    IRWLS_LAUNCH_CHECK();
    IRWLS_HIP(hipStreamSynchronize(g_stream));
    IRWLS_HIP(hipFree(probe));
    setstate(caller_state);
    return 0;
}

void irwls_dev_shutdown(void) {
    if (!g_init) return;
    IRWLS_HIP(hipStreamSynchronize(g_stream));
    scratch_release();
    cache_release();
    resolve_pool();
    IRWLS_HIP(hipStreamDestroy(g_stream));
    g_stream = nullptr;
    g_init = false;
}

void *irwls_dev_alloc(size_t bytes) {
    irwls_dev_init();
    if (bytes == 0) bytes = 8;
    size_t got = 0;
    void *p = cached_malloc(bytes, &got);
    irwls_dev_memset0(p, bytes);
    return p;
}

void irwls_dev_free(void *p) {
    if (!p) return;
    IRWLS_HIP(hipStreamSynchronize(stream()));
    cached_free(p);
}

// Small host-to-device copies go through a pinned staging ring: the caller's
// buffer is copied into the ring at once (so it may be reused on return) and
// the transfer is queued on the library stream without waiting, stream order
// keeping it ahead of the kernels that read it.  The ring restarts whenever
// the stream has drained (device-to-host copies and explicit
// synchronisations wait for it), or, when full, after a synchronisation.
static const size_t kStage = 16u << 20;
static const size_t kStageMax = 1u << 20;   // larger copies go directly and wait
static thread_local char *g_stage = nullptr;
static thread_local size_t g_stage_head = 0;
// Worker threads of a batched run wait for their stream through an event
// created with hipEventBlockingSync: the thread sleeps until the device
// signals instead of spinning, so many waiting workers leave the host cores
// to the ones with host work.  They also read back through a pinned buffer
// of their own (kBounce bytes) rather than the runtime's pageable path.
static thread_local hipEvent_t t_sync_event = nullptr;
static thread_local char *t_bounce = nullptr;
static const size_t kBounce = (size_t)1 << 20;

static thread_local double t_wait_s = 0.0;
static thread_local void *t_rb = nullptr;
static thread_local hipEvent_t t_rb_event = nullptr;

static double now_s() {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void wait_stream() {
    const double t0 = now_s();
    if (t_sync_event) {
        IRWLS_HIP(hipEventRecord(t_sync_event, stream()));
        IRWLS_HIP(hipEventSynchronize(t_sync_event));
    } else {
        IRWLS_HIP(hipStreamSynchronize(stream()));
    }
    t_wait_s += now_s() - t0;
}

static void stage_reset() { g_stage_head = 0; }

void irwls_dev_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    if (!g_stage) g_stage = (char *)pinned_get(kStage);
    if (bytes > kStageMax) {
        IRWLS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
        wait_stream();
        stage_reset();
        return;
    }
    if (g_stage_head + bytes > kStage) {
        wait_stream();
        stage_reset();
    }
    char *slot = g_stage + g_stage_head;
    memcpy(slot, src, bytes);
    IRWLS_HIP(hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, stream()));
    g_stage_head += (bytes + 255) & ~(size_t)255;
}

// Chunked uploads take their pinned buffers from the pool in one size.
static const size_t kChunkPool = (size_t)32 << 20;

void irwls_dev_h2d_chunked(void *dst, size_t total, size_t chunk,
                           void (*fill)(void *arg, void *buf, size_t off, size_t len), void *arg) {
    if (total == 0) return;
    if (chunk == 0 || chunk > total) chunk = total;
    const size_t cap = chunk <= kChunkPool ? kChunkPool : chunk;
    void *buf[2] = {pinned_get(cap), pinned_get(cap)};
    hipEvent_t ev[2];
    for (int b = 0; b < 2; b++) IRWLS_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
    bool busy[2] = {false, false};
    size_t k = 0;
    for (size_t off = 0; off < total; off += chunk, k++) {
        const int b = (int)(k & 1);
        const size_t len = (total - off < chunk) ? total - off : chunk;
        if (busy[b]) IRWLS_HIP(hipEventSynchronize(ev[b]));
        fill(arg, buf[b], off, len);
        IRWLS_HIP(hipMemcpyAsync((char *)dst + off, buf[b], len, hipMemcpyHostToDevice, stream()));
        IRWLS_HIP(hipEventRecord(ev[b], stream()));
        busy[b] = true;
    }
    for (int b = 0; b < 2; b++) {
        if (busy[b]) IRWLS_HIP(hipEventSynchronize(ev[b]));
        IRWLS_HIP(hipEventDestroy(ev[b]));
        pinned_put(buf[b], cap);
    }
}

void irwls_dev_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    if (t_bounce && bytes <= kBounce) {
        IRWLS_HIP(hipMemcpyAsync(t_bounce, src, bytes, hipMemcpyDeviceToHost, stream()));
        wait_stream();
        memcpy(dst, t_bounce, bytes);
    } else {
        IRWLS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
        wait_stream();
    }
    stage_reset();
}

// Two device-to-host copies with one wait for the stream: through the
// thread's pinned bounce buffer when both fit, else straight into dst.
static thread_local char *t_pair = nullptr;

void irwls_dev_d2h2(void *dst1, const void *src1, size_t b1, void *dst2, const void *src2, size_t b2) {
    if (b1 + b2 == 0) return;
    char *bb = t_bounce;
    if (!bb) {
        if (!t_pair) t_pair = (char *)pinned_get(kBounce);
        bb = t_pair;
    }
    const size_t o2 = (b1 + 255) & ~(size_t)255;
    if (o2 + b2 <= kBounce) {
        if (b1) IRWLS_HIP(hipMemcpyAsync(bb, src1, b1, hipMemcpyDeviceToHost, stream()));
        if (b2) IRWLS_HIP(hipMemcpyAsync(bb + o2, src2, b2, hipMemcpyDeviceToHost, stream()));
        wait_stream();
        if (b1) memcpy(dst1, bb, b1);
        if (b2) memcpy(dst2, bb + o2, b2);
    } else {
        if (b1) IRWLS_HIP(hipMemcpyAsync(dst1, src1, b1, hipMemcpyDeviceToHost, stream()));
        if (b2) IRWLS_HIP(hipMemcpyAsync(dst2, src2, b2, hipMemcpyDeviceToHost, stream()));
        wait_stream();
    }
    stage_reset();
}

}  // extern "C"

namespace irwls {

void readback_start(const void *src, size_t bytes) {
    if (bytes > kReadbackMax) {
        fprintf(stderr, "libirwls: readback of %zu bytes exceeds %zu\n", bytes, kReadbackMax);
        exit(2);
    }
    if (!t_rb) {
        t_rb = pinned_get(kReadbackMax);
        const unsigned fl = hipEventDisableTiming | (t_stream ? hipEventBlockingSync : 0u);
        IRWLS_HIP(hipEventCreateWithFlags(&t_rb_event, fl));
    }
    IRWLS_HIP(hipMemcpyAsync(t_rb, src, bytes, hipMemcpyDeviceToHost, stream()));
    IRWLS_HIP(hipEventRecord(t_rb_event, stream()));
}

const void *readback_wait() {
    const double t0 = now_s();
    IRWLS_HIP(hipEventSynchronize(t_rb_event));
    t_wait_s += now_s() - t0;
    return t_rb;
}

static void readback_release() {
    if (!t_rb) return;
    pinned_put(t_rb, kReadbackMax);
    t_rb = nullptr;
    IRWLS_HIP(hipEventDestroy(t_rb_event));
    t_rb_event = nullptr;
}

}  // namespace irwls

extern "C" {

void irwls_dev_d2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    IRWLS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
}

void irwls_dev_memset0(void *dst, size_t bytes) {
    if (bytes == 0) return;
    size_t work = (bytes + 7) / 8;
    int blocks = (int)((work + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, stream(),
                       (unsigned char *)dst, bytes);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_copy_kernel(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    if (((uintptr_t)dst & 15) || ((uintptr_t)src & 15) || (bytes & 15)) {
        fprintf(stderr, "libirwls: irwls_dev_copy_kernel needs 16-byte aligned pointers and size\n");
        exit(2);
    }
    const size_t n16 = bytes / 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, stream(), (const uint4 *)src,
                       (uint4 *)dst, n16);
    IRWLS_LAUNCH_CHECK();
}

void irwls_dev_sync(void) {
    wait_stream();
    stage_reset();
}

void irwls_dev_trim(void) {
    if (!g_init) return;
    IRWLS_HIP(hipStreamSynchronize(g_stream));
    scratch_release();
    cache_release();
}

void irwls_dev_timing_mask(unsigned mask) { g_timing_mask = mask; }

void irwls_dev_timing_stride(int stride) { g_stride = stride > 1 ? stride : 1; }

void irwls_dev_timing(int enable) {
    irwls_dev_init();
    resolve_pool();
    g_timing = enable != 0;
    memset(g_seen, 0, sizeof(g_seen));
    memset(g_launch, 0, sizeof(g_launch));
    memset(g_ms, 0, sizeof(g_ms));
    memset(g_flops, 0, sizeof(g_flops));
    memset(g_bytes, 0, sizeof(g_bytes));
}

void irwls_dev_timer_get(int kind, long *launches, double *ms, double *flops, double *bytes) {
    if (g_init) resolve_pool();
    if (kind < 0 || kind >= IRWLS_T_COUNT) {
        *launches = 0; *ms = 0; *flops = 0; *bytes = 0;
        return;
    }
    *launches = g_launch[kind];
    *ms = g_ms[kind];
    *flops = g_flops[kind];
    *bytes = g_bytes[kind];
}

void irwls_dev_release_thread_buffers(void) {
    irwls_dev_release_cholesky_buffers();
    irwls_dev_release_pcg_buffers();
    irwls_dev_release_mlp_buffers();
}

double irwls_dev_wait_seconds(int reset) {
    const double w = t_wait_s;
    if (reset) t_wait_s = 0.0;
    return w;
}

// Worker streams are created by the thread that starts the workers
// (irwls_dev_reserve_streams) and pooled, so no stream is ever created by a
// worker thread that ends while other threads still use the stream.
void irwls_dev_reserve_streams(int n) {
    irwls_dev_init();
    std::lock_guard<std::mutex> g(g_pmu);
    while ((int)g_streams.size() < n) {
        hipStream_t s = nullptr;
        IRWLS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        g_streams.push_back(s);
    }
}

void irwls_dev_thread_begin(void) {
    irwls_dev_init();
    if (t_stream) return;
    {
        std::lock_guard<std::mutex> g(g_pmu);
        if (!g_streams.empty()) {
            t_stream = g_streams.back();
            g_streams.pop_back();
        }
    }
    if (!t_stream) IRWLS_HIP(hipStreamCreateWithFlags(&t_stream, hipStreamNonBlocking));
    IRWLS_HIP(hipEventCreateWithFlags(&t_sync_event, hipEventBlockingSync | hipEventDisableTiming));
    t_bounce = (char *)pinned_get(kBounce);
}

void irwls_dev_thread_end(void) {
    if (!t_stream) return;
    IRWLS_HIP(hipStreamSynchronize(t_stream));
    irwls_dev_release_thread_buffers();
    scratch_release();
    readback_release();
    pinned_put(g_stage, kStage);
    g_stage = nullptr;
    g_stage_head = 0;
    pinned_put(t_bounce, kBounce);
    t_bounce = nullptr;
    IRWLS_HIP(hipEventDestroy(t_sync_event));
    t_sync_event = nullptr;
    std::lock_guard<std::mutex> g(g_pmu);
    for (hipStream_t s : g_streams)
        if (s == t_stream) {
            fprintf(stderr, "libirwls: worker stream returned to its pool twice\n");
            exit(2);
        }
    g_streams.push_back(t_stream);
    t_stream = nullptr;
}

}  // extern "C"
