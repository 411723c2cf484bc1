// C ABI of the GF(2^128) sumcheck engine: device memory, per-round handles
// and the full prover drivers (host loop around the HIP kernels).
//
// Each handle owns one HIP stream.  A round launches its kernels, copies the
// 16..48-byte round sums into pinned host memory and synchronises; the
// challenge for the next round then comes from the caller (handle API) or
// the transcript (drivers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bitz_hip.h"
#include "kernels.hpp"
#include "transcript.hpp"

using bitz::kBlock;
using bitz::Transcript;


namespace {

thread_local std::string g_err;

int32_t fail(int32_t code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(BITZ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define TRY(expr)                  \
    do {                           \
        int32_t rc_ = (expr);      \
        if (rc_ != BITZ_OK) return rc_; \
    } while (0)

inline f128 F(bitz_f128 x) { return f128{x.lo, x.hi}; }
inline bitz_f128 B(f128 x) { return bitz_f128{x.lo, x.hi}; }

// Device allocations are recycled through a process-wide cache keyed by
// size: a proof's scratch (tens of MiB to GiB) would otherwise pay hipMalloc
// and hipFree on every call.
struct DeviceCache {
    std::mutex mu;
    std::multimap<size_t, void*> free_blocks;
    size_t cached_bytes = 0;
    // Sizes above 64 KiB round up to a quarter-octave class (at most 25%
    // over), so that the blocks of requests of similar sizes are reused.
    static size_t size_class(size_t n) {
        if (n <= ((size_t)64 << 10)) return n;
        int lg = 63 - __builtin_clzll((unsigned long long)n);
        const size_t step = (size_t)1 << (lg - 2);
        return (n + step - 1) & ~(step - 1);
    }
    // A block of at least n bytes: one of n's own class when n is above
    // 64 KiB (a smaller request taking a larger class's block would leave
    // the next request of that class to a hipMalloc, which stalls every
    // thread's HIP calls for hundreds of microseconds), at most 2 n below;
    // *size receives its true size, which put() must be given back.
    void* get(size_t n, size_t* size) {
        n = size_class(n);
        {
            std::lock_guard<std::mutex> lock(mu);
            auto it = free_blocks.lower_bound(n);
            if (it != free_blocks.end() && (n > ((size_t)64 << 10) ? it->first == n : it->first <= 2 * n)) {
                void* p = it->second;
                *size = it->first;
                cached_bytes -= it->first;
                free_blocks.erase(it);
                return p;
            }
        }
        const auto t0 = std::chrono::steady_clock::now();
        void* p = nullptr;
        if (hipMalloc(&p, n) != hipSuccess) {
            // drop the cache and retry once
            release_all();
            if (hipMalloc(&p, n) != hipSuccess) return nullptr;
        }
        note_malloc(t0);
        *size = n;
        return p;
    }
    static void note_malloc(std::chrono::steady_clock::time_point t0);
    void put(void* p, size_t n) {
        std::lock_guard<std::mutex> lock(mu);
        free_blocks.emplace(n, p);
        cached_bytes += n;
    }
    void release_all() {
        std::lock_guard<std::mutex> lock(mu);
        (void)hipDeviceSynchronize();
        for (auto& kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
        cached_bytes = 0;
    }
} g_cache;

// A device allocation owned by a handle.
struct Dev {
    void* p = nullptr;
    size_t bytes = 0;
    Dev() = default;
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    ~Dev() { release(); }
    void release() {
        if (p) g_cache.put(p, bytes);
        p = nullptr;
        bytes = 0;
    }
    int32_t alloc(size_t n) {
        if (p) {
            // Replacing a buffer a queued kernel may still use: let the
            // device drain before the block returns to the cache.
            (void)hipDeviceSynchronize();
            release();
        }
        if (n == 0) n = 16;
        n = (n + 255) & ~(size_t)255;
        size_t got = 0;
        p = g_cache.get(n, &got);
        if (!p) return fail(BITZ_ERR_ALLOC, "hipMalloc of " + std::to_string(n) + " bytes failed");
        bytes = got;
        return BITZ_OK;
    }
    // At least n bytes, growing geometrically; the old block goes back to the
    // cache without a device drain, so the caller must know that nothing
    // queued still uses it.
    int32_t grow_idle(size_t n) {
        if (p && bytes >= n) return BITZ_OK;
        const size_t want = std::max<size_t>({n, 2 * bytes, (size_t)64 << 10});
        release();
        return alloc(want);
    }
    f128* f() const { return static_cast<f128*>(p); }
    uint32_t* u32() const { return static_cast<uint32_t*>(p); }
};

// Pinned host staging for round sums: fixed-size blocks recycled through a
// process-wide pool (hipHostMalloc per proof would cost more than a round).
struct PinnedPool {
    static constexpr size_t kBytes = 64 * 1024;
    std::mutex mu;
    std::vector<void*> free_blocks;
    void* get() {
        {
            std::lock_guard<std::mutex> lock(mu);
            if (!free_blocks.empty()) {
                void* p = free_blocks.back();
                free_blocks.pop_back();
                return p;
            }
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, kBytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        return p;
    }
    void put(void* p) {
        std::lock_guard<std::mutex> lock(mu);
        free_blocks.push_back(p);
    }
} g_pinned;

// Larger pinned staging: power-of-two classes from 64 KiB, recycled.
struct PinnedClassPool {
    std::mutex mu;
    std::multimap<size_t, void*> free_blocks;
    void* get(size_t n, size_t* size) {
        size_t c = PinnedPool::kBytes;
        while (c < n) c <<= 1;
        {
            std::lock_guard<std::mutex> lock(mu);
            auto it = free_blocks.find(c);
            if (it != free_blocks.end()) {
                void* p = it->second;
                free_blocks.erase(it);
                *size = c;
                return p;
            }
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
        *size = c;
        return p;
    }
    void put(void* p, size_t size) {
        std::lock_guard<std::mutex> lock(mu);
        free_blocks.emplace(size, p);
    }
} g_pinned_classes;

struct PinnedBuf {
    void* p = nullptr;
    size_t size = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) g_pinned_classes.put(p, size);
    }
    int32_t alloc(size_t n) {
        if (p && size >= n) return BITZ_OK;
        if (p) g_pinned_classes.put(p, size);
        p = g_pinned_classes.get(n, &size);
        if (!p) return fail(BITZ_ERR_ALLOC, "hipHostMalloc failed");
        return BITZ_OK;
    }
    char* c() const { return static_cast<char*>(p); }
};

struct Pinned {
    void* p = nullptr;
    Pinned() = default;
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
    ~Pinned() {
        if (p) g_pinned.put(p);
    }
    int32_t alloc(size_t n) {
        if (n > PinnedPool::kBytes) return fail(BITZ_ERR_ARG, "pinned staging request too large");
        if (!p) p = g_pinned.get();
        if (!p) return fail(BITZ_ERR_ALLOC, "hipHostMalloc failed");
        return BITZ_OK;
    }
};

unsigned grid_for(size_t work) {
    size_t g = (work + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (unsigned)g;
}

int32_t check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(BITZ_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return BITZ_OK;
}

// Copies n elements from the caller (host or device) into a fresh device buffer.
int32_t upload(Dev& d, const bitz_f128* src, size_t n, uint32_t flags, hipStream_t s) {
    TRY(d.alloc(n * sizeof(f128)));
    HIP_TRY(hipMemcpyAsync(d.p, src, n * sizeof(f128),
                           (flags & BITZ_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    return BITZ_OK;
}

int32_t download(bitz_f128* dst, const f128* src, size_t n, uint32_t flags, hipStream_t s) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(f128),
                           (flags & BITZ_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BITZ_OK;
}

// eq(b, r) over b in {0,1}^k into out (2^k entries); r is a device array.
// Up to 2^12 entries one workgroup doubles the table level by level; above,
// two half tables are built that way and multiplied entrywise.
// k_eq_rows' dynamic LDS (lane tables of 16 waves) is above the default
// limit: set once, from whichever thread gets here first.
static void eq_rows_attr() {
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_eq_rows),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem1024);
    });
}

int32_t eq_build(const f128* r_dev, int k, f128* out, Dev& scratch, hipStream_t s) {
    if (k <= bitz::kEqRowsMaxVars) {
        eq_rows_attr();
        hipLaunchKernelGGL(bitz::k_eq_rows, dim3(bitz::eq_rows_grid(k)), dim3(1024), bitz::kLaneSmem1024, s, r_dev, k,
                           out, static_cast<const f128*>(nullptr));
        return check_launch();
    }
    const int kl = k / 2, kh = k - kl;
    const size_t nl = (size_t)1 << kl, nh = (size_t)1 << kh;
    if (scratch.bytes < (nl + nh) * sizeof(f128)) TRY(scratch.alloc((nl + nh) * sizeof(f128)));
    f128* low = scratch.f();
    f128* high = low + nl;
    hipLaunchKernelGGL(bitz::k_eq_small2, dim3(2), dim3(kBlock), 0, s, r_dev, kl, low, kh, high);
    const size_t n = (size_t)1 << k;
    hipLaunchKernelGGL(bitz::k_eq_combine, dim3(grid_for(n)), dim3(kBlock), 0, s, low, kl, high, out, n);
    return check_launch();
}

int log2_exact(size_t n) {
    int k = 0;
    while (((size_t)1 << k) < n) ++k;
    return k;
}

// Streams are recycled through a pool as well: one per live handle, all at
// the default priority.
struct StreamPool {
    std::mutex mu;
    std::vector<hipStream_t> free_streams;
    hipStream_t get() {
        {
            std::lock_guard<std::mutex> lock(mu);
            if (!free_streams.empty()) {
                hipStream_t s = free_streams.back();
                free_streams.pop_back();
                return s;
            }
        }
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        return s;
    }
    void put(hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu);
        free_streams.push_back(s);
    }
} g_streams;

struct StreamOwner {
    hipStream_t s = nullptr;
    ~StreamOwner() {
        if (s) {
            (void)hipStreamSynchronize(s);
            g_streams.put(s);
        }
    }
    int32_t create() {
        s = g_streams.get();
        if (!s) return fail(BITZ_ERR_DEVICE, "hipStreamCreateWithFlags failed");
        return BITZ_OK;
    }
};

// Per-kernel timing: when enabled, launches on the sumcheck paths are
// bracketed by HIP events on their own stream; durations are read back when
// the round's sums are (the stream is synchronised there anyway).
struct KernelStat {
    double ms = 0, bytes = 0;
    uint64_t launches = 0;
};
struct Profiler {
    bool on = false;
    std::string only;  // when set, only launches of this name are timed
    std::mutex mu;
    std::map<std::string, KernelStat> stats;
} g_prof;

// Adds host time to a profiling entry (the engine's host phases).
static void prof_add_host(const char* name, std::chrono::steady_clock::time_point since, double bytes = 0) {
    if (!g_prof.on) return;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - since).count();
    std::lock_guard<std::mutex> lock(g_prof.mu);
    if (!g_prof.only.empty() && g_prof.only != name) return;
    KernelStat& k = g_prof.stats[name];
    k.ms += ms;
    k.bytes += bytes;
    k.launches += 1;
}

// Cache misses: the time of each hipMalloc the device cache had to make.
void DeviceCache::note_malloc(std::chrono::steady_clock::time_point t0) { prof_add_host("device_malloc", t0); }

// Events come from a pool so that a profiled proof creates none.
struct EventPool {
    std::mutex mu;
    std::vector<hipEvent_t> free_events;
    hipEvent_t get() {
        {
            std::lock_guard<std::mutex> lock(mu);
            if (!free_events.empty()) {
                hipEvent_t e = free_events.back();
                free_events.pop_back();
                return e;
            }
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void put(hipEvent_t e) {
        std::lock_guard<std::mutex> lock(mu);
        free_events.push_back(e);
    }
} g_events;

struct Timed {
    hipEvent_t start = nullptr, stop = nullptr;
    const char* name = nullptr;
    double bytes = 0;
    uint64_t launches = 1;  // back-to-back launches between the two events
    Timed() = default;
    Timed(const Timed&) = delete;
    Timed& operator=(const Timed&) = delete;
    ~Timed() {
        if (start) g_events.put(start);
        if (stop) g_events.put(stop);
    }
    // Records the start event if profiling is on.
    void begin(hipStream_t s, const char* kernel, double algo_bytes) {
        name = nullptr;
        if (!g_prof.on) return;
        if (!g_prof.only.empty() && g_prof.only != kernel) return;
        if (!start) start = g_events.get();
        if (!stop) stop = g_events.get();
        if (!start || !stop) return;
        name = kernel;
        bytes = algo_bytes;
        (void)hipEventRecord(start, s);
    }
    void end(hipStream_t s) {
        if (name) (void)hipEventRecord(stop, s);
    }
    // Call after the stream has been synchronised.
    void collect() {
        if (!name) return;
        float ms = 0;
        if (hipEventElapsedTime(&ms, start, stop) == hipSuccess) {
            std::lock_guard<std::mutex> lock(g_prof.mu);
            KernelStat& k = g_prof.stats[name];
            k.ms += ms;
            k.bytes += bytes;
            k.launches += launches;
        }
        name = nullptr;
    }
};

}  // namespace

// ====================================================== static scratch ===
// The allocation-free queued inner prover
// (bitz_prove_inner_sumcheck_with_challenges_noalloc) works in this array:
// every GPU the library runs on holds its own copy from module load until the
// process ends.  Host code never uses the array's name as an address: each
// device's address comes from hipGetSymbolAddress, resolved once per device
// with that device current (bitz_init, or the entry point's first call on
// the device).  Layout for 2^n entries, n <= kNoAllocMaxVars: P (D, h folded
// once: 2 x 2^(n-1) entries) | Q (2 x 2^(n-2)) | n challenge tables | the n
// challenges | the round sums ((n + 2) x 8 words).
namespace {
constexpr int kNoAllocMaxVars = 20;
constexpr size_t kNoAllocQOff = ((size_t)1 << kNoAllocMaxVars) * sizeof(f128);   // 16 MiB
constexpr size_t kNoAllocTabOff = kNoAllocQOff + (kNoAllocQOff >> 1);             // 24 MiB
constexpr size_t kNoAllocChalOff = kNoAllocTabOff + (size_t)kNoAllocMaxVars * kConstTable4Entries * sizeof(uint4);
constexpr size_t kNoAllocAccOff = kNoAllocChalOff + 512;
constexpr size_t kNoAllocScratchBytes = kNoAllocAccOff + 1024;
static_assert(kNoAllocMaxVars * sizeof(f128) <= 512, "challenges fit their slot");
static_assert((kNoAllocMaxVars + 2) * 32 <= 1024, "round sums fit their slot");
constexpr int kMaxDevices = 64;
struct ScratchAddrs {
    std::mutex mu;
    void* ptr[kMaxDevices] = {};
} g_scratch;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// This device's copy of `scratch` (the current device's), resolved on first use.
static int32_t scratch_address(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(BITZ_ERR_ARG, "device index out of range for the static scratch");
    std::lock_guard<std::mutex> lock(g_scratch.mu);
    if (!g_scratch.ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BITZ_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch.ptr[dev] = p;
    }
    *out = g_scratch.ptr[dev];
    return BITZ_OK;
}

// =============================================================== runtime ===
extern "C" int32_t bitz_init(int32_t device) {
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipFree(nullptr));
    void* p = nullptr;
    return scratch_address(&p);
}

extern "C" const char* bitz_last_error(void) { return g_err.c_str(); }

namespace bitz {
// Error reporting for the host-only translation units (sha256_r1cs.cpp).
int32_t report_error(int32_t code, const std::string& msg) { return fail(code, msg); }
}  // namespace bitz

// Caller allocations come from the same size-keyed cache as the engine's
// scratch: hipFree synchronizes the whole device, which would stall every
// other stream (concurrent requests) at each release.
struct Outstanding {
    std::mutex mu;
    std::map<void*, size_t> sizes;
} g_outstanding;

extern "C" int32_t bitz_device_alloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(BITZ_ERR_ARG, "null pointer");
    size_t n = 0;
    void* p = g_cache.get(bytes ? bytes : 16, &n);
    if (!p) return fail(BITZ_ERR_ALLOC, "hipMalloc failed");
    {
        std::lock_guard<std::mutex> lock(g_outstanding.mu);
        g_outstanding.sizes[p] = n;
    }
    *ptr = p;
    return BITZ_OK;
}
extern "C" int32_t bitz_device_free(void* ptr) {
    if (!ptr) return BITZ_OK;
    size_t n = 0;
    {
        std::lock_guard<std::mutex> lock(g_outstanding.mu);
        auto it = g_outstanding.sizes.find(ptr);
        if (it != g_outstanding.sizes.end()) {
            n = it->second;
            g_outstanding.sizes.erase(it);
        }
    }
    if (n) {
        g_cache.put(ptr, n);
        return BITZ_OK;
    }
    HIP_TRY(hipFree(ptr));
    return BITZ_OK;
}
extern "C" int32_t bitz_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BITZ_OK;
}
extern "C" int32_t bitz_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BITZ_OK;
}
extern "C" int32_t bitz_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return BITZ_OK;
}
extern "C" int32_t bitz_release_cache(void) {
    g_cache.release_all();
    return BITZ_OK;
}

extern "C" int32_t bitz_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return BITZ_OK;
}

extern "C" int32_t bitz_profile_enable(int32_t on) {
    std::lock_guard<std::mutex> lock(g_prof.mu);
    g_prof.stats.clear();
    g_prof.on = on != 0;
    return BITZ_OK;
}

extern "C" int32_t bitz_profile_only(const char* kernel) {
    std::lock_guard<std::mutex> lock(g_prof.mu);
    g_prof.only = kernel ? kernel : "";
    return BITZ_OK;
}

extern "C" int32_t bitz_profile_read(const char* kernel, double* total_ms, uint64_t* launches,
                                     double* algo_bytes) {
    if (!kernel || !total_ms || !launches || !algo_bytes) return fail(BITZ_ERR_ARG, "null pointer");
    std::lock_guard<std::mutex> lock(g_prof.mu);
    auto it = g_prof.stats.find(kernel);
    if (it == g_prof.stats.end()) {
        *total_ms = 0;
        *launches = 0;
        *algo_bytes = 0;
    } else {
        *total_ms = it->second.ms;
        *launches = it->second.launches;
        *algo_bytes = it->second.bytes;
    }
    return BITZ_OK;
}

// =========================================================== transcripts ===
extern "C" bitz_transcript* bitz_transcript_callbacks(void* ctx, bitz_absorb_fn absorb,
                                                      bitz_squeeze_fn squeeze) {
    if (!absorb || !squeeze) return nullptr;
    auto* t = new bitz_transcript;
    t->t.type = Transcript::Type::kCallbacks;
    t->t.cb_ctx = ctx;
    t->t.cb_absorb = absorb;
    t->t.cb_squeeze = squeeze;
    return t;
}
extern "C" int32_t bitz_transcript_set_grind(bitz_transcript* t, bitz_grind_fn grind) {
    if (!t) return fail(BITZ_ERR_ARG, "null pointer");
    if (t->t.type != Transcript::Type::kCallbacks) return fail(BITZ_ERR_ARG, "only callback transcripts take a grind hook");
    t->t.cb_grind = grind;
    return BITZ_OK;
}
extern "C" int32_t bitz_transcript_grind(bitz_transcript* t, const uint8_t* label, size_t len, uint32_t bits) {
    if (!t || (len && !label)) return fail(BITZ_ERR_ARG, "null pointer");
    if (bits > bitz::kMaxGrindingBits) return fail(BITZ_ERR_ARG, "grinding difficulty exceeds 32 bits");
    t->t.grind(label, len, bits);
    return BITZ_OK;
}
extern "C" int32_t bitz_pow_find(const uint8_t seed[16], uint32_t bits, uint64_t* nonce) {
    if (!seed || !nonce) return fail(BITZ_ERR_ARG, "null pointer");
    if (bits > bitz::kMaxGrindingBits) return fail(BITZ_ERR_ARG, "grinding difficulty exceeds 32 bits");
    *nonce = bitz::pow_find(seed, bits);
    return BITZ_OK;
}
extern "C" int32_t bitz_pow_valid(const uint8_t seed[16], uint64_t nonce, uint32_t bits) {
    return seed && bitz::pow_valid(seed, nonce, bits) ? 1 : 0;
}
extern "C" uint32_t bitz_grinding_bits(uint32_t security_bits, uint64_t coefficient) {
    return bitz::grinding_bits(security_bits, coefficient);
}
extern "C" bitz_transcript* bitz_transcript_keccak(const uint8_t* domain, size_t len) {
    if (len && !domain) return nullptr;
    auto* t = new bitz_transcript;
    t->t.type = Transcript::Type::kKeccak;
    bitz::keccak_start(t->t.keccak, domain, len);
    return t;
}
extern "C" int32_t bitz_keccak_hash(const uint8_t* msg, size_t len, int32_t rate, uint8_t pad, uint8_t* out,
                                    size_t out_len, uint32_t flags) {
    if ((len && !msg) || (out_len && !out)) return fail(BITZ_ERR_ARG, "null pointer");
    if (rate <= 0 || rate > 200 || rate % 8) return fail(BITZ_ERR_ARG, "the rate is a multiple of 8 bytes below 200");
    if (!(flags & BITZ_DEVICE)) {
        bitz::keccak_hash(msg, len, rate, pad, out, out_len);
        return BITZ_OK;
    }
    // The same permutation run by one device thread (the device code's check).
    StreamOwner so;
    TRY(so.create());
    Dev dm, dout;
    TRY(dm.alloc(std::max<size_t>(len, 1)));
    TRY(dout.alloc(std::max<size_t>(out_len, 1)));
    if (len) HIP_TRY(hipMemcpyAsync(dm.p, msg, len, hipMemcpyHostToDevice, so.s));
    hipLaunchKernelGGL(bitz::k_keccak_hash, dim3(1), dim3(64), 0, so.s, static_cast<const uint8_t*>(dm.p), len,
                       (int)rate, pad, static_cast<uint8_t*>(dout.p), out_len);
    TRY(check_launch());
    if (out_len) HIP_TRY(hipMemcpyAsync(out, dout.p, out_len, hipMemcpyDeviceToHost, so.s));
    HIP_TRY(hipStreamSynchronize(so.s));
    return BITZ_OK;
}
extern "C" bitz_transcript* bitz_transcript_seeded(uint64_t seed) {
    auto* t = new bitz_transcript;
    t->t.type = Transcript::Type::kSeeded;
    t->t.seed = seed;
    return t;
}
extern "C" bitz_transcript* bitz_transcript_duplex(const uint8_t* domain, size_t len) {
    auto* t = new bitz_transcript;
    t->t.type = Transcript::Type::kDuplex;
    bitz::sha256(domain, len, t->t.state);
    return t;
}
extern "C" void bitz_transcript_free(bitz_transcript* t) { delete t; }
extern "C" int32_t bitz_transcript_absorb(bitz_transcript* t, int32_t kind, const uint8_t* bytes,
                                          size_t len) {
    if (!t || (len && !bytes)) return fail(BITZ_ERR_ARG, "null transcript or message");
    t->t.absorb(kind, bytes, len);
    return BITZ_OK;
}
extern "C" int32_t bitz_transcript_squeeze(bitz_transcript* t, bitz_f128* out) {
    if (!t || !out) return fail(BITZ_ERR_ARG, "null transcript or output");
    *out = B(t->t.squeeze());
    return BITZ_OK;
}
extern "C" size_t bitz_transcript_narg(const bitz_transcript* t, uint8_t* out, size_t cap) {
    if (!t) return 0;
    if (out) memcpy(out, t->t.narg.data(), std::min(cap, t->t.narg.size()));
    return t->t.narg.size();
}

// ============================================================ primitives ===
extern "C" int32_t bitz_mul_batch(const bitz_f128* a, const bitz_f128* b, bitz_f128* out, size_t n,
                                  uint32_t flags) {
    if (n && (!a || !b || !out)) return fail(BITZ_ERR_ARG, "null pointer");
    if (n == 0) return BITZ_OK;
    Dev da, db, dout;
    StreamOwner so;
    TRY(so.create());
    TRY(upload(da, a, n, flags, so.s));
    TRY(upload(db, b, n, flags, so.s));
    TRY(dout.alloc(n * sizeof(f128)));
    const uint32_t mode = flags & (BITZ_MUL_LANE | BITZ_MUL_LANE2 | BITZ_MUL_CONST | BITZ_MUL_ROWS | BITZ_MUL_MFMA);
    // One event pair around the launch (with profiling on): "mul_comb",
    // "mul_lane", "mul_lane2", "mul_const", "mul_rows" or "mul_mfma", with
    // 32 B read and 16 B written per product as its algorithmic bytes.
    static const char* const kNames[] = {"mul_comb", "mul_lane", "mul_lane2", "mul_const", "mul_rows", "mul_mfma"};
    const int which = mode == 0 ? 0 : mode == BITZ_MUL_LANE ? 1 : mode == BITZ_MUL_LANE2 ? 2
                    : mode == BITZ_MUL_CONST ? 3 : mode == BITZ_MUL_ROWS ? 4 : mode == BITZ_MUL_MFMA ? 5 : -1;
    if (which < 0) return fail(BITZ_ERR_ARG, "one product mode at a time");
    Timed tm;
    tm.begin(so.s, kNames[which], (which == 3 || which == 5 ? 32.0 : 48.0) * (double)n);
    if (mode == 0) {
        hipLaunchKernelGGL(bitz::k_mul, dim3(grid_for(n)), dim3(kBlock), 0, so.s, da.f(), db.f(), dout.f(), n);
    } else if (mode == BITZ_MUL_LANE || mode == BITZ_MUL_LANE2) {
        static bool attr = false;
        if (!attr) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_mul_lane),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
            attr = true;
        }
        hipLaunchKernelGGL(bitz::k_mul_lane, dim3(grid_for(n)), dim3(kBlock), bitz::kLaneSmem256, so.s, da.f(),
                           db.f(), dout.f(), n, mode == BITZ_MUL_LANE2 ? 1 : 0);
    } else if (mode == BITZ_MUL_CONST) {
        hipLaunchKernelGGL(bitz::k_mul_const, dim3(grid_for(n)), dim3(kBlock), 0, so.s, da.f(), db.f(), dout.f(), n);
    } else if (mode == BITZ_MUL_ROWS) {
        hipLaunchKernelGGL(bitz::k_mul_rows, dim3(grid_for(16 * n)), dim3(kBlock), 0, so.s, da.f(), db.f(), dout.f(), n);
    } else {
        hipLaunchKernelGGL(bitz::k_mul_mfma, dim3(grid_for(n)), dim3(kBlock), 0, so.s, da.f(), db.f(), dout.f(), n);
    }
    tm.end(so.s);
    TRY(check_launch());
    TRY(download(out, dout.f(), n, flags, so.s));  // synchronises the stream
    tm.collect();
    return BITZ_OK;
}

extern "C" int32_t bitz_eq_table(const bitz_f128* r, uint32_t k, bitz_f128* out, uint32_t flags) {
    if (!out || (k && !r)) return fail(BITZ_ERR_ARG, "null pointer");
    if (k > 34) return fail(BITZ_ERR_ARG, "eq table too large");
    Dev dr, dout, scratch;
    StreamOwner so;
    TRY(so.create());
    TRY(upload(dr, r, k, flags, so.s));
    const size_t n = (size_t)1 << k;
    TRY(dout.alloc(n * sizeof(f128)));
    TRY(eq_build(dr.f(), (int)k, dout.f(), scratch, so.s));
    return download(out, dout.f(), n, flags, so.s);
}

// Folds the `len` entries at `in` (a device table, read in place and never
// written) by r[0 .. m), two variables a pass while two or more remain, into
// the scratch tables s0 (len / 2 entries) and s1 (len / 4); returns the table
// that holds the len >> m results.  The first pass is timed as
// "mle_fold_pass" (its algorithmic bytes: every input read, every output
// written once).
static const f128* mle_fold_passes(const f128* in, size_t len, const bitz_f128* r, uint32_t m, f128* s0, f128* s1,
                                   hipStream_t s, Timed& first) {
    const f128* cur = in;
    f128* bufs[2] = {s0, s1};
    int w = 0;
    for (uint32_t i = 0; i < m;) {
        f128* out = bufs[w];
        w ^= 1;
        const bool two = m - i >= 2;
        const size_t out_len = two ? len / 4 : len / 2;
        if (i == 0) first.begin(s, "mle_fold_pass", 16.0 * (double)(len + out_len));
        if (two)
            hipLaunchKernelGGL(bitz::k_fold2_t, dim3(grid_for(out_len)), dim3(kBlock), 0, s, cur, out, out_len, F(r[i]),
                               F(r[i + 1]));
        else
            hipLaunchKernelGGL(bitz::k_fold_t, dim3(grid_for(out_len)), dim3(kBlock), 0, s, cur, out, out_len, F(r[i]));
        if (i == 0) first.end(s);
        i += two ? 2 : 1;
        len = out_len;
        cur = out;
    }
    return cur;
}

extern "C" int32_t bitz_mle_fold(bitz_f128* evals, uint32_t* num_vars, const bitz_f128* r, uint32_t m,
                                 uint32_t flags) {
    if (!evals || !num_vars || (m && !r)) return fail(BITZ_ERR_ARG, "null pointer");
    // mle.rs:136-142: too many challenges is rejected before mutating.
    if (m > *num_vars) return fail(BITZ_ERR_TOO_MANY, "more challenges than variables");
    if (m == 0) return BITZ_OK;
    const bool dev = (flags & BITZ_DEVICE) != 0;
    Dev a, b, c;
    StreamOwner so;
    TRY(so.create());
    const size_t n = (size_t)1 << *num_vars;
    const f128* in = reinterpret_cast<const f128*>(evals);
    if (!dev) {
        TRY(upload(c, evals, n, flags, so.s));
        in = c.f();
    }
    TRY(a.alloc(std::max<size_t>(n / 2, 1) * sizeof(f128)));
    TRY(b.alloc(std::max<size_t>(n / 4, 1) * sizeof(f128)));
    std::vector<bitz_f128> rr(m);
    HIP_TRY(hipMemcpy(rr.data(), r, m * sizeof(f128), dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
    Timed tm;
    const f128* res = mle_fold_passes(in, n, rr.data(), m, a.f(), b.f(), so.s, tm);
    TRY(check_launch());
    TRY(download(evals, res, n >> m, flags, so.s));  // in place: the first n >> m entries
    tm.collect();
    *num_vars -= m;
    return BITZ_OK;
}

extern "C" int32_t bitz_mle_evaluate(const bitz_f128* evals, uint32_t num_vars, const bitz_f128* point,
                                     uint32_t point_len, bitz_f128* out, uint32_t flags) {
    if (!evals || !out || (point_len && !point)) return fail(BITZ_ERR_ARG, "null pointer");
    if (point_len != num_vars) return fail(BITZ_ERR_ARG, "point width differs from num_vars");
    const bool dev = (flags & BITZ_DEVICE) != 0;
    Dev a, b, c;
    StreamOwner so;
    TRY(so.create());
    const size_t n = (size_t)1 << num_vars;
    const f128* in = reinterpret_cast<const f128*>(evals);
    if (!dev) {
        TRY(upload(c, evals, n, flags, so.s));
        in = c.f();
    }
    TRY(a.alloc(std::max<size_t>(n / 2, 1) * sizeof(f128)));
    TRY(b.alloc(std::max<size_t>(n / 4, 1) * sizeof(f128)));
    std::vector<bitz_f128> rr(point_len);
    if (point_len)
        HIP_TRY(hipMemcpy(rr.data(), point, point_len * sizeof(f128), dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
    Timed tm;
    const f128* res = mle_fold_passes(in, n, rr.data(), point_len, a.f(), b.f(), so.s, tm);
    TRY(check_launch());
    bitz_f128 v;
    HIP_TRY(hipMemcpyAsync(&v, res, sizeof v, hipMemcpyDeviceToHost, so.s));
    HIP_TRY(hipStreamSynchronize(so.s));
    tm.collect();
    if (dev)
        HIP_TRY(hipMemcpy(out, &v, sizeof v, hipMemcpyHostToDevice));
    else
        *out = v;
    return BITZ_OK;
}

// ======================================================= outer sumcheck ===
// Rounds whose tables hold at most 2^kTailLog elements run in the
// single-workgroup tail kernel (pipelined prover only).
constexpr int kTailLog = 10;
static_assert(kTailLog + 1 <= bitz::kTailMaxLog, "tail tables must fit the kernel's LDS budget");

struct bitz_outer {
    int n = 0;
    size_t N = 0, len = 0, low_len = 0, high_len = 0;
    int round = 0;
    bool started = false, finished = false;
    Dev in, P, Q;      // inputs (owned copy when given on host), ping-pong scratch
    const f128* inputs[3] = {nullptr, nullptr, nullptr};
    const f128* cur[3] = {nullptr, nullptr, nullptr};
    int where = 0;     // 0 = inputs, 1 = P, 2 = Q
    size_t capP = 0, capQ = 0;
    // eq factors: eq_in[] are the caller's device tables (or an uploaded
    // copy); the general kernels fold an owned copy (eqL/eqH ping-pong).
    const f128* eq_in[2] = {nullptr, nullptr};
    bool factors_owned = false;
    Dev eqL[2], eqH[2];
    int curL = 0, curH = 0;
    unsigned eq_bl = 1, eq_bh = 1;  // factor-check workgroups over eq_low, eq_high
    Dev acc;
    Pinned host;
    Timed tm;

    // Eq-factored mode: the factors are eq tables of a point tau.  The
    // device checks them (k_eq_check) while the first round runs; the host
    // reads the verdict with the first round sums.
    bool eqmode = false;   // eq-factored kernels in use
    bool eq_valid = false; // the host has read and accepted the check
    int ell = 0, m = 0;
    std::vector<f128> tau, tau_inv;
    std::vector<f128> scale_big, scale_tail;  // divide out the factors' constants
    std::vector<f128> kpre[2], ksuf[2];       // prefix / suffix products of (1 + tau_i) per factor
    Dev htabs4, rtabs4;
    struct Run {  // variables [a, b) read from factor `table` (0 = eq_low, 1 = eq_high)
        int table, a, b;
    };
    struct Geo {
        size_t threads, iters;
        Run la, lb;
        int la_bits;
        int h_start;       // first iteration variable
        size_t htab_off;   // first 4-bit H(i) table (multi-iteration rounds)
    };
    std::vector<Geo> geo;
    bitz::HtabDesc hdesc{};
    size_t htab_total = 0;  // 4-bit H(i) tables of the multi-iteration rounds
    f128 S{1, 0}, sigma{0, 0};
    f128 q[3] = {{0, 0}, {0, 0}, {0, 0}};  // current round's q0, q1, q2

    // acc layout (32-bit words): round k's sums at 12 k; the factor check at
    // eq_off() (n + 2 probes, then 8 words per check workgroup); the
    // pipelined prover's final A(r), B(r), C(r) at fin_off().
    size_t eq_off() const { return 12 * (size_t)(n + 2); }
    size_t fin_off() const { return eq_off() + 4 * (size_t)(n + 2) + 8 * (size_t)(eq_bl + eq_bh); }
    size_t acc_words() const { return fin_off() + 12; }
    const f128* table_ptr(int t) const { return eq_in[t]; }
    int table_base(int t) const { return t == 0 ? 0 : ell; }

    int32_t read_words(size_t words) {
        HIP_TRY(hipMemcpyAsync(host.p, acc.p, words * 4, hipMemcpyDeviceToHost, so.s));
        HIP_TRY(hipStreamSynchronize(so.s));
        tm.collect();
        return BITZ_OK;
    }
    int32_t read_slot(int slot, bitz_f128* out3) {
        HIP_TRY(hipMemcpyAsync(host.p, acc.u32() + 12 * slot, 48, hipMemcpyDeviceToHost, so.s));
        HIP_TRY(hipStreamSynchronize(so.s));
        memcpy(out3, host.p, 48);
        tm.collect();
        return BITZ_OK;
    }
    // A batch member's sums live in the batch's shared buffers.
    bool batch_member = false;
    int tail_log = kTailLog;  // a batch member's grid-wide rounds end at n - tail_log
    const uint32_t* host_ext = nullptr;
    const uint32_t* host_words() const {
        return host_ext ? host_ext : static_cast<const uint32_t*>(host.p);
    }
    // Back to the round-0 state (for a second pipelined attempt).
    int32_t rewind() {
        for (int t = 0; t < 3; ++t) cur[t] = inputs[t];
        where = 0;
        len = N;
        round = 0;
        S = f128{1, 0};
        sigma = f128{0, 0};
        HIP_TRY(hipMemsetAsync(acc.p, 0, acc.bytes, so.s));
        return BITZ_OK;
    }
    StreamOwner so;  // last: destroyed (and synchronised) before the buffers
};

static int32_t outer_check(uint32_t low_vars, uint32_t high_vars, uint32_t az_vars, uint32_t bz_vars,
                           uint32_t cz_vars) {
    // sumcheck.rs:258-269, before any transcript mutation.
    if (bz_vars != az_vars || cz_vars != az_vars)
        return fail(BITZ_ERR_PRODUCT_DIMS, "product tables have different num_vars");
    if ((uint64_t)low_vars + high_vars != az_vars)
        return fail(BITZ_ERR_EQUALITY_DIMS, "eq factor num_vars do not sum to the product num_vars");
    if (az_vars > 34) return fail(BITZ_ERR_ARG, "num_vars too large");
    return BITZ_OK;
}

// Montgomery batch inversion of the nonzero entries (zeros stay zero).
static void batch_inverse(std::vector<f128>& v) {
    std::vector<f128> pre(v.size());
    f128 acc = hostf::one();
    for (size_t i = 0; i < v.size(); ++i) {
        pre[i] = acc;
        if (!hostf::is_zero(v[i])) acc = hostf::mul(acc, v[i]);
    }
    f128 inv = hostf::inv(acc);
    for (size_t i = v.size(); i-- > 0;) {
        if (hostf::is_zero(v[i])) continue;
        const f128 vi = v[i];
        v[i] = hostf::mul(inv, pre[i]);
        inv = hostf::mul(inv, vi);
    }
}

static f128 f128_at(const uint32_t* w) {
    f128 x;
    memcpy(&x, w, sizeof x);
    return x;
}


// Plans the eq-factored mode (round geometry, H(i) tables) and allocates
// its buffers.  Returns false (general kernels) when the factors are too
// large to check cheaply.  The check itself is launched by
// outer_launch_checks or, in the pipelined prover, by its prologue; nothing
// reads back here: the verdict arrives with the first round sums.
static bool outer_setup_eq(bitz_outer* h, int low_vars, int high_vars, bool batch_member = false,
                           int thread_bits = bitz::kThreadBits) {
    const int n = h->n;
    if (n == 0) return false;
    const size_t nl = (size_t)1 << low_vars, nh = (size_t)1 << high_vars;
    if (nl + nh > ((size_t)1 << 18)) return false;
    h->ell = low_vars;
    h->m = high_vars;
    const int ell = low_vars;
    bitz::HtabDesc desc{};
    h->geo.resize(n);
    size_t htab_total = 0;
    for (int k = 0; k < n; ++k) {
        const int pbits = n - k - 1;  // log2 pairs of round k
        // log2 threads: one wave of lane-table workgroups across the chip;
        // larger rounds iterate.
        const int tbits = std::min(pbits, thread_bits);
        auto& g = h->geo[k];
        g.threads = (size_t)1 << tbits;
        g.iters = (size_t)1 << (pbits - tbits);
        const int a = k + 1, b = a + tbits;
        if (a < ell && ell < b) {  // the thread variables cross from eq_low into eq_high
            g.la = {0, a, ell};
            g.lb = {1, ell, b};
        } else {
            const int t = a < ell ? 0 : 1;
            const int lab = std::min(tbits, 9);
            g.la = {t, a, a + lab};
            g.lb = {t, a + lab, b};
        }
        g.la_bits = g.la.b - g.la.a;
        g.h_start = b;
        g.htab_off = 0;
        if (g.iters > 1) {
            desc.first[desc.rounds] = (uint32_t)htab_total;
            desc.start[desc.rounds] = b;
            desc.rounds += 1;
            g.htab_off = htab_total;
            htab_total += g.iters;
        }
    }
    desc.first[desc.rounds] = (uint32_t)htab_total;
    h->hdesc = desc;
    h->htab_total = htab_total;
    if (!batch_member &&
        (h->rtabs4.alloc((size_t)std::max(n, 1) * kConstTable4Entries * sizeof(uint4)) ||
         h->htabs4.alloc(std::max<size_t>(htab_total, 1) * kConstTable4Entries * sizeof(uint4))))
        return false;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_outer_r0<false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kR0Smem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_outer_r0<true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kR0Smem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_outer_fold<false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kFoldSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_outer_fold<true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kFoldSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_outer_tail),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bitz::outer_stage_smem(bitz::kTailMaxLog - 1));
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_outer_chunk),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bitz::outer_stage_smem(bitz::kChunkMaxLog - 1));
        attr_set = true;
    }
    return true;
}

static unsigned eq_check_blocks(size_t entries, unsigned block) {
    return (unsigned)std::max<size_t>(1, (entries + block - 1) / block);
}

// The factor check and the H(i) tables, for the per-round prover.
static int32_t outer_launch_checks(bitz_outer* h) {
    hipStream_t s = h->so.s;
    hipLaunchKernelGGL(bitz::k_eq_check, dim3(h->eq_bl + h->eq_bh), dim3(1024), 0, s, h->eq_in[0], h->ell,
                       h->eq_in[1], h->m, h->acc.u32() + h->eq_off(), h->eq_bl);
    if (h->htab_total)
        hipLaunchKernelGGL(bitz::k_htabs4, dim3((unsigned)h->htab_total), dim3(1024), 0, s, h->eq_in[0], h->ell,
                           h->eq_in[1], h->hdesc, static_cast<uint4*>(h->htabs4.p));
    return check_launch();
}

// Round schedule of the pipelined eq-factored prover: rounds [0, kc) as
// grid-wide kernels, [kc, k0) in the chunked kernel, [k0, n) in the tail.
struct PipeSchedule {
    int kc, k0;
};
static PipeSchedule pipe_schedule(int n, int tail_log = kTailLog) {
    const int k0 = std::max(0, n - tail_log);
    const int kc = std::max(1, n - 1 - bitz::kChunkFromBits);  // rounds before kc run grid-wide
    if (k0 - kc >= 1) return {kc, k0};
    return {k0, k0};
}

// Constant carried by run_weight's product over the variables [a, b)
// (kernels.hip): eq_low's part times eq_high's part, each the product of
// 1 + tau_i over that factor's variables outside the run.
static f128 run_constant(const bitz_outer* h, int a, int b) {
    const int ell = h->ell;
    auto kap = [&](int t, int x, int y) {
        const int lo = t == 0 ? 0 : ell;
        return hostf::mul(h->kpre[t][x - lo], h->ksuf[t][y - lo]);
    };
    if (b <= ell) return kap(0, a, b);
    if (a >= ell) return kap(1, a, b);
    return hostf::mul(kap(0, a, ell), kap(1, ell, b));
}

// Constants of the weights of an LDS stage (lds_stage in kernels.hip):
// round k = k_first + l weighs its pairs by E(y 2^l) W_hi(g), where E runs
// over [a0, b) = [k_first + 1, k_first + 1 + pl0) and W_hi over [b, n), and
// the l zero variables of y 2^l contribute prod_{a0 <= i <= k} (1 + tau_i).
static void stage_constants(const bitz_outer* h, int k_first, int s, int pl0, f128* kap) {
    const int a0 = k_first + 1, b = a0 + pl0;
    f128 c = hostf::mul(run_constant(h, a0, b), run_constant(h, b, h->n));
    for (int l = 0; l < s; ++l) {
        const int k = k_first + l;
        if (l > 0) c = hostf::mul(c, hostf::add(hostf::one(), h->tau[k]));
        kap[k] = c;
    }
}

// Reads the device check (eq region of the host copy of acc) and, when the
// factors are eq tables, recovers tau and the constants the kernels' weights
// carry.  For a factor T with variables V and a run R of them, T[x << s]
// = c_R eq(tau_R, x) with c_R = prod_{i in V \ R} (1 + tau_i), since
// T[0] = prod_V (1 + tau_i) and T[2^i] / T[0] = tau_i / (1 + tau_i).
static bool outer_eq_constants(bitz_outer* h, const std::vector<f128>& tau);

static bool outer_validate_eq(bitz_outer* h, const uint32_t* region) {
    const int n = h->n, ell = h->ell;
    const uint32_t* probes = region;
    const uint32_t* rec = region + 4 * (n + 2);
    f128 sum[2] = {{0, 0}, {0, 0}};
    for (unsigned b = 0; b < h->eq_bl + h->eq_bh; ++b) {
        if (rec[8 * b]) return false;
        const f128 x = f128_at(rec + 8 * b + 4);
        sum[b >= h->eq_bl] = hostf::add(sum[b >= h->eq_bl], x);
    }
    if (!hostf::eq(sum[0], hostf::one()) || !hostf::eq(sum[1], hostf::one())) return false;
    std::vector<f128> den(n), num(n);
    for (int t = 0; t < 2; ++t) {
        const int base = t == 0 ? 0 : ell, bits = t == 0 ? ell : h->m;
        const uint32_t* P = probes + 4 * (t == 0 ? 0 : ell + 1);
        const f128 t0 = f128_at(P);
        if (hostf::is_zero(t0)) return false;
        for (int i = 0; i < bits; ++i) {
            num[base + i] = f128_at(P + 4 * (1 + i));
            den[base + i] = hostf::add(t0, num[base + i]);
            if (hostf::is_zero(den[base + i])) return false;
        }
    }
    batch_inverse(den);
    std::vector<f128> tau(n);
    for (int i = 0; i < n; ++i) tau[i] = hostf::mul(num[i], den[i]);
    return outer_eq_constants(h, tau);
}

// The eq-factored prover's constants from the point tau: the run constants of
// every round's weights, their inverses (the round scales) and 1 / tau_k.
// False when some 1 + tau_i is zero (the factors then vanish somewhere).
static bool outer_eq_constants(bitz_outer* h, const std::vector<f128>& tau) {
    const int n = h->n, ell = h->ell;
    std::vector<f128> one_tau(n);
    for (int i = 0; i < n; ++i) {
        one_tau[i] = hostf::add(hostf::one(), tau[i]);
        if (hostf::is_zero(one_tau[i])) return false;
    }
    // Per factor t with variables [lo, hi): pre[t][j] = prod_{lo <= i < lo + j}
    // (1 + tau_i), suf[t][j] = prod_{lo + j <= i < hi}.  For a run [a, b)
    // of its variables, c = pre[t][a - lo] * suf[t][b - lo].
    std::vector<f128>(&pre)[2] = h->kpre;
    std::vector<f128>(&suf)[2] = h->ksuf;
    for (int t = 0; t < 2; ++t) {
        const int lo = t == 0 ? 0 : ell, hi = t == 0 ? ell : n, w = hi - lo;
        pre[t].assign(w + 1, hostf::one());
        suf[t].assign(w + 1, hostf::one());
        for (int j = 0; j < w; ++j) pre[t][j + 1] = hostf::mul(pre[t][j], one_tau[lo + j]);
        for (int j = w; j-- > 0;) suf[t][j] = hostf::mul(suf[t][j + 1], one_tau[lo + j]);
    }
    auto kappa = [&](int t, int a, int b) {
        const int lo = t == 0 ? 0 : ell, hi = t == 0 ? ell : n;
        a = std::max(lo, std::min(a, hi));
        b = std::max(a, std::min(b, hi));
        return hostf::mul(pre[t][a - lo], suf[t][b - lo]);
    };
    std::vector<f128> kap(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const auto& g = h->geo[k];
        f128 c = hostf::mul(kappa(g.la.table, g.la.a, g.la.b), kappa(g.lb.table, g.lb.a, g.lb.b));
        if (g.iters > 1) {
            if (g.h_start < ell)
                c = hostf::mul(c, hostf::mul(kappa(0, g.h_start, ell), kappa(1, ell, n)));
            else
                c = hostf::mul(c, kappa(1, g.h_start, n));
        }
        kap[k] = c;
        kap[n + k] = hostf::one();
    }
    // Rounds of the pipelined prover's LDS stages (chunk and tail).
    h->tau = tau;
    const PipeSchedule ps = pipe_schedule(n, h->tail_log);
    if (ps.kc < ps.k0) stage_constants(h, ps.kc, ps.k0 - ps.kc, bitz::kChunkMaxLog - 2, kap.data() + n);
    stage_constants(h, ps.k0, n - ps.k0, n - ps.k0 - 1, kap.data() + n);
    batch_inverse(kap);
    h->scale_big.assign(kap.begin(), kap.begin() + n);
    h->scale_tail.assign(kap.begin() + n, kap.end());
    h->tau_inv = tau;
    batch_inverse(h->tau_inv);
    h->eq_valid = true;
    return true;
}

// Launches the eq-factored round-k kernel into slot k: round 0 through
// k_outer_r0, the fold rounds through k_outer_fold (lane-table products,
// folding with the 4-bit table rtab4 of r_{k-1}).
// With `timed` false the launch is left to a span timer of the caller.
static int32_t outer_eq_launch(bitz_outer* h, int k, const uint4* rtab4, bool q1,
                               f128* const out[3], Timed* tm = nullptr, bool timed = true) {
    if (!tm) tm = &h->tm;
    const auto& g = h->geo[k];
    uint32_t* slot = h->acc.u32() + 12 * k;
    hipStream_t s = h->so.s;
    const size_t pairs = g.threads * g.iters;
    if (timed) tm->begin(s, k == 0 ? "outer_round0" : "outer_fold_next", k == 0 ? 96.0 * pairs : 288.0 * pairs);
    const uint4* h4 = static_cast<const uint4*>(h->htabs4.p) + g.htab_off * kConstTable4Entries;
    const f128* la = h->table_ptr(g.la.table);
    const f128* lb = h->table_ptr(g.lb.table);
    const int la_shift = g.la.a - h->table_base(g.la.table), lb_shift = g.lb.a - h->table_base(g.lb.table);
    if (k == 0) {
        const unsigned grid = (unsigned)((g.threads + bitz::kR0Block - 1) / bitz::kR0Block);
#define OUTER_R0_ARGS \
    h->cur[0], h->cur[1], h->cur[2], g.threads, g.iters, la, g.la_bits, la_shift, lb, lb_shift, h4, slot
        if (q1)
            hipLaunchKernelGGL((bitz::k_outer_r0<true>), dim3(grid), dim3(bitz::kR0Block), bitz::kR0Smem, s,
                               OUTER_R0_ARGS);
        else
            hipLaunchKernelGGL((bitz::k_outer_r0<false>), dim3(grid), dim3(bitz::kR0Block), bitz::kR0Smem, s,
                               OUTER_R0_ARGS);
#undef OUTER_R0_ARGS
    } else if (g.iters > 1) {
        const unsigned grid = (unsigned)((g.threads + bitz::kFoldBlock - 1) / bitz::kFoldBlock);
#define OUTER_FOLD_ARGS                                                                                  \
    h->cur[0], h->cur[1], h->cur[2], out[0], out[1], out[2], g.threads, g.iters, rtab4, la, g.la_bits, la_shift, \
        lb, lb_shift, h4, slot
        if (q1)
            hipLaunchKernelGGL((bitz::k_outer_fold<true>), dim3(grid), dim3(bitz::kFoldBlock), bitz::kFoldSmem, s,
                               OUTER_FOLD_ARGS);
        else
            hipLaunchKernelGGL((bitz::k_outer_fold<false>), dim3(grid), dim3(bitz::kFoldBlock), bitz::kFoldSmem, s,
                               OUTER_FOLD_ARGS);
#undef OUTER_FOLD_ARGS
    } else {
        // Single iteration: the lane-table fold kernel in workgroups of at
        // most 256 threads (whole waves) so small rounds spread over the CUs.
        const unsigned block = (unsigned)std::min<size_t>(256, (g.threads + 63) / 64 * 64);
        const unsigned grid = (unsigned)((g.threads + block - 1) / block);
        const size_t smem = bitz::fold_smem(block);
#define OUTER_FOLD1_ARGS                                                                                   \
    h->cur[0], h->cur[1], h->cur[2], out[0], out[1], out[2], g.threads, g.iters, rtab4, la, g.la_bits, la_shift, \
        lb, lb_shift, h4, slot
        if (q1)
            hipLaunchKernelGGL((bitz::k_outer_fold<true>), dim3(grid), dim3(block), smem, s, OUTER_FOLD1_ARGS);
        else
            hipLaunchKernelGGL((bitz::k_outer_fold<false>), dim3(grid), dim3(block), smem, s, OUTER_FOLD1_ARGS);
#undef OUTER_FOLD1_ARGS
    }
    if (timed) tm->end(s);
    return check_launch();
}

// Round k's [c0, c2, c3] from the kernel's sums (times the round's scale):
// q(0), q(inf) and q(1) when computed, else q(1) recovered from
// sigma_k = (1 + tau_k) q(0) + tau_k q(1).  g_k(X) = S_k (1 + tau_k + X) q_k(X).
static void outer_eq_coeffs(bitz_outer* h, int k, const bitz_f128* raw, bool have_q1, const f128& scale,
                            bitz_f128* co) {
    const f128 tau = h->tau[k];
    const f128 one_tau = hostf::add(hostf::one(), tau);
    const f128 q0 = hostf::mul(F(raw[0]), scale), q2 = hostf::mul(F(raw[1]), scale);
    f128 q_at1;
    if (have_q1)
        q_at1 = hostf::mul(F(raw[2]), scale);
    else
        q_at1 = hostf::mul(hostf::add(h->sigma, hostf::mul(one_tau, q0)), h->tau_inv[k]);
    const f128 q1 = hostf::add(hostf::add(q_at1, q0), q2);
    h->q[0] = q0;
    h->q[1] = q1;
    h->q[2] = q2;
    const f128 S = h->S;
    co[0] = B(hostf::mul(S, hostf::mul(one_tau, q0)));
    co[1] = B(hostf::mul(S, hostf::add(hostf::mul(one_tau, q2), q1)));
    co[2] = B(hostf::mul(S, q2));
}

// After round k's challenge r: sigma_{k+1} = q_k(r), S_{k+1} = S_k eq(tau_k, r).
static void outer_eq_advance(bitz_outer* h, int k, f128 r) {
    h->sigma = hostf::add(h->q[0], hostf::mul(r, hostf::add(h->q[1], hostf::mul(r, h->q[2]))));
    const f128 e = f128{1 ^ h->tau[k].lo ^ r.lo, h->tau[k].hi ^ r.hi};
    h->S = hostf::mul(h->S, e);
}

// The general kernels fold the eq factors: give them an owned, writable
// copy (the caller's tables are borrowed) and the half-size ping-pong tables.
static int32_t outer_own_factors(bitz_outer* h) {
    hipStream_t s = h->so.s;
    if (!h->factors_owned) {
        TRY(h->eqL[0].alloc(h->low_len * sizeof(f128)));
        TRY(h->eqH[0].alloc(h->high_len * sizeof(f128)));
        HIP_TRY(hipMemcpyAsync(h->eqL[0].p, h->eq_in[0], h->low_len * sizeof(f128), hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(h->eqH[0].p, h->eq_in[1], h->high_len * sizeof(f128), hipMemcpyDeviceToDevice, s));
        h->factors_owned = true;
    }
    if (!h->eqL[1].p) TRY(h->eqL[1].alloc(std::max<size_t>(h->low_len / 2, 1) * sizeof(f128)));
    if (!h->eqH[1].p) TRY(h->eqH[1].alloc(std::max<size_t>(h->high_len / 2, 1) * sizeof(f128)));
    return BITZ_OK;
}

// launch_checks: launch the factor check and H(i) tables now (per-round
// prover); the pipelined prover folds them into its prologue instead.
static int32_t outer_begin_impl(bitz_outer** out, const bitz_f128* eq_low, uint32_t low_vars,
                                const bitz_f128* eq_high, uint32_t high_vars, const bitz_f128* az,
                                uint32_t az_vars, const bitz_f128* bz, uint32_t bz_vars, const bitz_f128* cz,
                                uint32_t cz_vars, uint32_t flags, bool launch_checks) {
    if (!out || !eq_low || !eq_high || !az || !bz || !cz) return fail(BITZ_ERR_ARG, "null pointer");
    TRY(outer_check(low_vars, high_vars, az_vars, bz_vars, cz_vars));
    auto* h = new bitz_outer;
    auto cleanup = [&](int32_t rc) {
        delete h;
        return rc;
    };
    int32_t rc;
    if ((rc = h->so.create())) return cleanup(rc);
    hipStream_t s = h->so.s;
    h->n = (int)az_vars;
    const size_t N = (size_t)1 << az_vars;
    h->N = N;
    h->len = N;
    h->low_len = (size_t)1 << low_vars;
    h->high_len = (size_t)1 << high_vars;
    if (flags & BITZ_DEVICE) {
        h->inputs[0] = reinterpret_cast<const f128*>(az);
        h->inputs[1] = reinterpret_cast<const f128*>(bz);
        h->inputs[2] = reinterpret_cast<const f128*>(cz);
    } else {
        if ((rc = h->in.alloc(3 * N * sizeof(f128)))) return cleanup(rc);
        const bitz_f128* src[3] = {az, bz, cz};
        for (int t = 0; t < 3; ++t) {
            if (hipMemcpyAsync(h->in.f() + t * N, src[t], N * sizeof(f128), hipMemcpyHostToDevice, s) !=
                hipSuccess)
                return cleanup(fail(BITZ_ERR_DEVICE, "upload of product tables failed"));
            h->inputs[t] = h->in.f() + t * N;
        }
    }
    for (int t = 0; t < 3; ++t) h->cur[t] = h->inputs[t];
    h->capP = std::max<size_t>(N / 2, 1);
    h->capQ = std::max<size_t>(N / 4, 1);
    if ((rc = h->P.alloc(3 * h->capP * sizeof(f128)))) return cleanup(rc);
    if ((rc = h->Q.alloc(3 * h->capQ * sizeof(f128)))) return cleanup(rc);
    if (flags & BITZ_DEVICE) {
        h->eq_in[0] = reinterpret_cast<const f128*>(eq_low);
        h->eq_in[1] = reinterpret_cast<const f128*>(eq_high);
    } else {
        if ((rc = upload(h->eqL[0], eq_low, h->low_len, flags, s))) return cleanup(rc);
        if ((rc = upload(h->eqH[0], eq_high, h->high_len, flags, s))) return cleanup(rc);
        h->eq_in[0] = h->eqL[0].f();
        h->eq_in[1] = h->eqH[0].f();
        h->factors_owned = true;
    }
    h->eq_bl = eq_check_blocks(h->low_len, 1024);
    h->eq_bh = eq_check_blocks(h->high_len, 1024);
    if ((rc = h->acc.alloc(h->acc_words() * sizeof(uint32_t)))) return cleanup(rc);
    if (h->acc_words() * sizeof(uint32_t) > PinnedPool::kBytes / 2)
        return cleanup(fail(BITZ_ERR_ARG, "eq factors too large"));
    // The pipelined prover's prologue clears what it needs.
    if (launch_checks && hipMemsetAsync(h->acc.p, 0, h->acc.bytes, s) != hipSuccess)
        return cleanup(fail(BITZ_ERR_DEVICE, "memset failed"));
    if ((rc = h->host.alloc(h->acc_words() * sizeof(uint32_t)))) return cleanup(rc);
    h->eqmode = outer_setup_eq(h, (int)low_vars, (int)high_vars);
    if (h->eqmode && launch_checks && outer_launch_checks(h)) h->eqmode = false;
    if (!h->eqmode && (rc = outer_own_factors(h))) return cleanup(rc);
    (void)hipGetLastError();  // a failed probe must not leak into later checks
    *out = h;
    return BITZ_OK;
}

extern "C" int32_t bitz_outer_begin(bitz_outer** out, const bitz_f128* eq_low, uint32_t low_vars,
                                    const bitz_f128* eq_high, uint32_t high_vars, const bitz_f128* az,
                                    uint32_t az_vars, const bitz_f128* bz, uint32_t bz_vars,
                                    const bitz_f128* cz, uint32_t cz_vars, uint32_t flags) {
    return outer_begin_impl(out, eq_low, low_vars, eq_high, high_vars, az, az_vars, bz, bz_vars, cz, cz_vars,
                            flags, true);
}

// Reads the device check of the factors if the host has not yet.
static int32_t outer_confirm_eq(bitz_outer* h) {
    if (!h->eqmode || h->eq_valid) return BITZ_OK;
    TRY(h->read_words(h->fin_off()));
    if (!outer_validate_eq(h, h->host_words() + h->eq_off())) h->eqmode = false;
    return BITZ_OK;
}

extern "C" int32_t bitz_outer_eq_mode(bitz_outer* h) {
    if (!h) return 0;
    if (outer_confirm_eq(h)) return 0;
    return h->eqmode ? 1 : 0;
}

static int32_t outer_generic_round0(bitz_outer* h) {
    const size_t pairs = h->len / 2;
    h->tm.begin(h->so.s, "outer_round0", 48.0 * (double)h->len);
    hipLaunchKernelGGL(bitz::k_outer_round0, dim3(grid_for(pairs)), dim3(kBlock), 0, h->so.s, h->cur[0],
                       h->cur[1], h->cur[2], pairs, h->eqL[h->curL].f(), log2_exact(h->low_len),
                       h->eqH[h->curH].f(), h->acc.u32());
    h->tm.end(h->so.s);
    return check_launch();
}

extern "C" int32_t bitz_outer_round0(bitz_outer* h, bitz_f128* co) {
    if (!h || !co) return fail(BITZ_ERR_ARG, "null pointer");
    if (h->started) return fail(BITZ_ERR_STATE, "round0 called twice");
    h->started = true;
    if (h->n == 0) {
        memset(co, 0, 48);
        return BITZ_OK;
    }
    if (h->eqmode) {
        // q(1) is computed: the claim the caller holds is not trusted here.
        // The factor check is read with the first sums.
        TRY(outer_eq_launch(h, 0, nullptr, true, nullptr));
        bitz_f128 raw[3];
        if (h->eq_valid) {
            TRY(h->read_slot(0, raw));
        } else {
            TRY(h->read_words(h->fin_off()));
            memcpy(raw, h->host.p, 48);
            if (!outer_validate_eq(h, h->host_words() + h->eq_off())) h->eqmode = false;
        }
        if (h->eqmode) {
            outer_eq_coeffs(h, 0, raw, true, h->scale_big[0], co);
            return BITZ_OK;
        }
        // Not eq tables of a point: the general kernels take over.
        HIP_TRY(hipMemsetAsync(h->acc.p, 0, 48, h->so.s));
        TRY(outer_own_factors(h));
    }
    TRY(outer_generic_round0(h));
    return h->read_slot(0, co);
}

extern "C" int32_t bitz_outer_fold_next(bitz_outer* h, bitz_f128 rb, bitz_f128* co) {
    if (!h || !co) return fail(BITZ_ERR_ARG, "null pointer");
    if (!h->started || h->round >= h->n) return fail(BITZ_ERR_STATE, "no round left to fold");
    hipStream_t s = h->so.s;
    const f128 r = F(rb);
    const size_t next_len = h->len / 2;
    // Output buffers: P after the inputs or Q, Q after P.
    Dev& dst = (h->where == 1) ? h->Q : h->P;
    const size_t cap = (h->where == 1) ? h->capQ : h->capP;
    f128* out[3] = {dst.f(), dst.f() + cap, dst.f() + 2 * cap};
    uint32_t* slot = h->acc.u32() + 12 * (h->round + 1);
    bool has_next = false;
    if (h->eqmode) {
        const int k = h->round;  // r is round k's challenge
        outer_eq_advance(h, k, r);
        if (k + 1 < h->n) {
            uint4* rtab4 = static_cast<uint4*>(h->rtabs4.p);
            hipLaunchKernelGGL(bitz::k_const_table4_1, dim3(1), dim3(1024), 0, s, r, rtab4);
            const bool q1 = hostf::is_zero(h->tau[k + 1]);
            TRY(outer_eq_launch(h, k + 1, rtab4, q1, out));
            for (int t = 0; t < 3; ++t) h->cur[t] = out[t];
            h->where = (h->where == 1) ? 2 : 1;
            h->len = next_len;
            h->round += 1;
            bitz_f128 raw[3];
            TRY(h->read_slot(h->round, raw));
            outer_eq_coeffs(h, h->round, raw, q1, h->scale_big[h->round], co);
            return BITZ_OK;
        }
        hipLaunchKernelGGL(bitz::k_fold3, dim3(1), dim3(kBlock), 0, s, h->cur[0], h->cur[1], h->cur[2], out[0],
                           out[1], out[2], (size_t)1, r);
        TRY(check_launch());
        for (int t = 0; t < 3; ++t) h->cur[t] = out[t];
        h->where = (h->where == 1) ? 2 : 1;
        h->len = next_len;
        h->round += 1;
        memset(co, 0, 48);
        return BITZ_OK;
    }
    if (h->low_len > 1) {
        // Phase 1 (sumcheck.rs:294-323): fold eq_low, then the products.
        const size_t nl = h->low_len / 2;
        hipLaunchKernelGGL(bitz::k_fold, dim3(grid_for(nl)), dim3(kBlock), 0, s, h->eqL[h->curL].f(),
                           h->eqL[1 - h->curL].f(), nl, r);
        h->curL ^= 1;
        h->low_len = nl;
        if (next_len > 1) {
            h->tm.begin(s, "outer_fold_next", 72.0 * (double)h->len);
            hipLaunchKernelGGL(bitz::k_outer_fold_next, dim3(grid_for(next_len / 2)), dim3(kBlock), 0, s,
                               h->cur[0], h->cur[1], h->cur[2], out[0], out[1], out[2], next_len / 2, r,
                               h->eqL[h->curL].f(), log2_exact(h->low_len), h->eqH[h->curH].f(), slot);
            h->tm.end(s);
            has_next = true;
        } else {
            hipLaunchKernelGGL(bitz::k_fold3, dim3(1), dim3(kBlock), 0, s, h->cur[0], h->cur[1], h->cur[2],
                               out[0], out[1], out[2], (size_t)1, r);
        }
    } else {
        // Phase 2 (sumcheck.rs:328-362): fold eq_high with the products.
        const size_t nh = h->high_len / 2;
        hipLaunchKernelGGL(bitz::k_fold, dim3(grid_for(nh)), dim3(kBlock), 0, s, h->eqH[h->curH].f(),
                           h->eqH[1 - h->curH].f(), nh, r);
        h->curH ^= 1;
        h->high_len = nh;
        if (nh == 1) {
            hipLaunchKernelGGL(bitz::k_fold3, dim3(1), dim3(kBlock), 0, s, h->cur[0], h->cur[1], h->cur[2],
                               out[0], out[1], out[2], (size_t)1, r);
        } else {
            h->tm.begin(s, "outer_fold_next", 72.0 * (double)h->len);
            hipLaunchKernelGGL(bitz::k_outer_fold_next, dim3(grid_for(nh / 2)), dim3(kBlock), 0, s,
                               h->cur[0], h->cur[1], h->cur[2], out[0], out[1], out[2], nh / 2, r,
                               h->eqL[h->curL].f(), 0, h->eqH[h->curH].f(), slot);
            h->tm.end(s);
            has_next = true;
        }
    }
    TRY(check_launch());
    for (int t = 0; t < 3; ++t) h->cur[t] = out[t];
    h->where = (h->where == 1) ? 2 : 1;
    h->len = next_len;
    h->round += 1;
    if (has_next) return h->read_slot(h->round, co);
    memset(co, 0, 48);
    return BITZ_OK;
}

extern "C" int32_t bitz_outer_finish(bitz_outer* h, bitz_f128* claims) {
    if (!h || !claims) return fail(BITZ_ERR_ARG, "null pointer");
    if (h->round != h->n) return fail(BITZ_ERR_STATE, "rounds remain");
    for (int t = 0; t < 3; ++t)
        HIP_TRY(hipMemcpyAsync(claims + t, h->cur[t], sizeof(f128), hipMemcpyDeviceToHost, h->so.s));
    HIP_TRY(hipStreamSynchronize(h->so.s));
    h->finished = true;
    return BITZ_OK;
}

extern "C" int32_t bitz_outer_tables(bitz_outer* h, bitz_f128* az, bitz_f128* bz, bitz_f128* cz,
                                     size_t* len) {
    if (!h || !len) return fail(BITZ_ERR_ARG, "null pointer");
    *len = h->len;
    bitz_f128* dst[3] = {az, bz, cz};
    for (int t = 0; t < 3; ++t)
        if (dst[t])
            HIP_TRY(hipMemcpyAsync(dst[t], h->cur[t], h->len * sizeof(f128), hipMemcpyDeviceToHost, h->so.s));
    HIP_TRY(hipStreamSynchronize(h->so.s));
    return BITZ_OK;
}

extern "C" void bitz_outer_free(bitz_outer* h) {
    if (h) {
        (void)hipStreamSynchronize(h->so.s);
        delete h;
    }
}

extern "C" int32_t bitz_outer_claim(const bitz_f128* eq_low, uint32_t low_vars, const bitz_f128* eq_high,
                                    uint32_t high_vars, const bitz_f128* az, const bitz_f128* bz,
                                    const bitz_f128* cz, uint32_t num_vars, uint32_t flags, bitz_f128* out) {
    if (!eq_low || !eq_high || !az || !bz || !cz || !out) return fail(BITZ_ERR_ARG, "null pointer");
    TRY(outer_check(low_vars, high_vars, num_vars, num_vars, num_vars));
    Dev in, lo, hi, acc;
    StreamOwner so;
    TRY(so.create());
    const size_t n = (size_t)1 << num_vars;
    const f128* t[3];
    if (flags & BITZ_DEVICE) {
        t[0] = reinterpret_cast<const f128*>(az);
        t[1] = reinterpret_cast<const f128*>(bz);
        t[2] = reinterpret_cast<const f128*>(cz);
    } else {
        TRY(in.alloc(3 * n * sizeof(f128)));
        const bitz_f128* src[3] = {az, bz, cz};
        for (int k = 0; k < 3; ++k) {
            HIP_TRY(hipMemcpyAsync(in.f() + k * n, src[k], n * sizeof(f128), hipMemcpyHostToDevice, so.s));
            t[k] = in.f() + k * n;
        }
    }
    TRY(upload(lo, eq_low, (size_t)1 << low_vars, flags, so.s));
    TRY(upload(hi, eq_high, (size_t)1 << high_vars, flags, so.s));
    TRY(acc.alloc(16));
    HIP_TRY(hipMemsetAsync(acc.p, 0, 16, so.s));
    hipLaunchKernelGGL(bitz::k_outer_claim, dim3(grid_for(n)), dim3(kBlock), 0, so.s, t[0], t[1], t[2], n, lo.f(),
                       (int)low_vars, hi.f(), acc.u32());
    TRY(check_launch());
    bitz_f128 v;
    HIP_TRY(hipMemcpyAsync(&v, acc.p, 16, hipMemcpyDeviceToHost, so.s));
    HIP_TRY(hipStreamSynchronize(so.s));
    *out = v;
    return BITZ_OK;
}

// sumcheck.rs:760-829: c1 = claim - c0 - (c0 + c2 + ... + cD), which in
// characteristic 2 is claim + c2 + ... + cD; absorb [c0..cD], squeeze r,
// claim <- g(r).
static f128 recover_round(Transcript& t, f128& claim, const bitz_f128* without_linear, int degree,
                          bitz_f128* round_out, bitz_f128* point_out) {
    f128 c[4];
    c[0] = F(without_linear[0]);
    f128 c1 = claim;
    for (int i = 2; i <= degree; ++i) {
        c[i] = F(without_linear[i - 1]);
        c1 = hostf::add(c1, c[i]);
    }
    c[1] = c1;
    t.absorb_elems(0, c, (size_t)degree + 1);
    const f128 r = t.squeeze();
    claim = hostf::horner(c, degree + 1, r);
    for (int i = 0; i <= degree; ++i) round_out[i] = B(c[i]);
    *point_out = B(r);
    return r;
}

extern "C" int32_t bitz_prove_outer_sumcheck(bitz_transcript* t, bitz_f128 initial_claim,
                                             const bitz_f128* eq_low, uint32_t low_vars,
                                             const bitz_f128* eq_high, uint32_t high_vars,
                                             const bitz_f128* az, uint32_t az_vars, const bitz_f128* bz,
                                             uint32_t bz_vars, const bitz_f128* cz, uint32_t cz_vars,
                                             uint32_t flags, bitz_f128* round_polys,
                                             bitz_f128* eval_points, bitz_f128* final_claim,
                                             bitz_f128* claims) {
    if (!t || !final_claim || !claims || (az_vars && (!round_polys || !eval_points)))
        return fail(BITZ_ERR_ARG, "null pointer");
    TRY(outer_check(low_vars, high_vars, az_vars, bz_vars, cz_vars));
    bitz_outer* h = nullptr;
    TRY(bitz_outer_begin(&h, eq_low, low_vars, eq_high, high_vars, az, az_vars, bz, bz_vars, cz, cz_vars,
                         flags));
    int32_t rc = BITZ_OK;
    bitz_f128 co[3];
    f128 claim = F(initial_claim);
    if ((rc = bitz_outer_round0(h, co))) goto done;
    for (uint32_t k = 0; k < az_vars; ++k) {
        const f128 r = recover_round(t->t, claim, co, 3, round_polys + 4 * k, eval_points + k);
        if ((rc = bitz_outer_fold_next(h, B(r), co))) goto done;
    }
    if ((rc = bitz_outer_finish(h, claims))) goto done;
    {
        f128 cl[3] = {F(claims[0]), F(claims[1]), F(claims[2])};
        t->t.absorb_elems(0, cl, 3);  // sumcheck.rs:372
    }
    *final_claim = B(claim);
done:
    bitz_outer_free(h);
    return rc;
}

// Launches the prologue: clearing of the round-sum slots, factor check,
// H(i) tables and the challenges' constant tables (the challenges travel
// as kernel arguments).
static int32_t outer_pipeline_prologue(bitz_outer* h, const bitz_f128* challenges) {
    hipStream_t s = h->so.s;
    const int n = h->n;
    if (n > bitz::kMaxRounds) return fail(BITZ_ERR_ARG, "num_vars too large");
    bitz::PrologueArgs pa{};
    pa.L = h->eq_in[0];
    pa.H = h->eq_in[1];
    pa.ell = h->ell;
    pa.m = h->m;
    pa.eq_blocks_l = h->eq_bl;
    pa.eq_blocks = h->eq_bl + h->eq_bh;
    pa.eq_out = h->acc.u32() + h->eq_off();
    pa.zero = h->acc.u32();
    pa.zero_words = (unsigned)h->eq_off();
    for (int k = 0; k < n; ++k) pa.chal[k] = F(challenges[k]);
    pa.h4_blocks = (unsigned)h->htab_total;
    pa.hdesc = h->hdesc;
    pa.htabs4 = static_cast<uint4*>(h->htabs4.p);
    pa.t4_blocks = (unsigned)n;
    pa.rtabs4 = static_cast<uint4*>(h->rtabs4.p);
    const unsigned grid = pa.eq_blocks + pa.h4_blocks + pa.t4_blocks;
    hipLaunchKernelGGL(bitz::k_prologue, dim3(grid), dim3(1024), 0, s, pa);
    return check_launch();
}

// Enqueues a whole eq-factored proof after its prologue (the challenges'
// constant tables are on the device), then reads every round's sums, the
// factor check and the final evaluations at once.  q1mask selects the
// rounds that compute q(1).
static int32_t outer_pipeline_eq(bitz_outer* h, uint64_t q1mask, std::vector<Timed>& timers) {
    hipStream_t s = h->so.s;
    const int n = h->n;
    const PipeSchedule ps = pipe_schedule(n);
    const uint4* rtabs4 = static_cast<const uint4*>(h->rtabs4.p);
    auto next_out = [&](f128* (&out)[3]) {
        Dev& dst = (h->where == 1) ? h->Q : h->P;
        const size_t cap = (h->where == 1) ? h->capQ : h->capP;
        out[0] = dst.f();
        out[1] = dst.f() + cap;
        out[2] = dst.f() + 2 * cap;
    };
    auto advance = [&](f128* const (&out)[3], size_t new_len) {
        for (int t = 0; t < 3; ++t) h->cur[t] = out[t];
        h->where = (h->where == 1) ? 2 : 1;
        h->len = new_len;
    };
    // The grid-wide fold rounds run back to back: one event pair spans them
    // (two event records per proof instead of two per launch).
    double fold_bytes = 0;
    for (int k = 1; k < ps.kc; ++k) fold_bytes += 288.0 * h->geo[k].threads * h->geo[k].iters;
    for (int k = 0; k < ps.kc; ++k) {
        f128* out[3] = {nullptr, nullptr, nullptr};
        if (k > 0) next_out(out);
        if (k == 1) {
            timers[1].begin(s, "outer_fold_next", fold_bytes);
            timers[1].launches = (uint64_t)(ps.kc - 1);
        }
        TRY(outer_eq_launch(h, k, k > 0 ? rtabs4 + (size_t)(k - 1) * kConstTable4Entries : nullptr,
                            (q1mask >> k) & 1, k > 0 ? out : nullptr, &timers[k], k == 0));
        if (k == ps.kc - 1 && k > 0) timers[1].end(s);
        if (k > 0) advance(out, h->len / 2);
    }
    if (ps.kc < ps.k0) {
        const int cs = bitz::kChunkMaxLog, steps = ps.k0 - ps.kc;
        f128* out[3];
        next_out(out);
        bitz::LdsStage ca{};
        ca.a = h->cur[0];
        ca.b = h->cur[1];
        ca.c = h->cur[2];
        ca.ao = out[0];
        ca.bo = out[1];
        ca.co = out[2];
        ca.rtabs4 = rtabs4;
        ca.L = h->eq_in[0];
        ca.H = h->eq_in[1];
        ca.ell = h->ell;
        ca.n = n;
        ca.k_first = ps.kc;
        ca.s = steps;
        ca.lg_in = cs;
        ca.fold_first = 1;
        ca.q1mask = q1mask;
        ca.acc = h->acc.u32();
        const unsigned chunks = (unsigned)(h->len >> cs);
        timers[ps.kc].begin(s, "outer_chunk", 0.0);
        hipLaunchKernelGGL(bitz::k_outer_chunk, dim3(chunks), dim3(bitz::kOuterStageBlock), bitz::outer_stage_smem(cs - 1), s,
                           ca);
        timers[ps.kc].end(s);
        TRY(check_launch());
        advance(out, h->len >> steps);
    }
    bitz::LdsStage ta{};
    ta.a = h->cur[0];
    ta.b = h->cur[1];
    ta.c = h->cur[2];
    ta.rtabs4 = rtabs4;
    ta.L = h->eq_in[0];
    ta.H = h->eq_in[1];
    ta.ell = h->ell;
    ta.n = n;
    ta.k_first = ps.k0;
    ta.s = n - ps.k0;
    ta.fold_first = ps.k0 > 0;
    ta.lg_in = ps.k0 > 0 ? n - ps.k0 + 1 : n;
    ta.final_stage = 1;
    ta.store_sums = 1;
    ta.q1mask = q1mask;
    ta.acc = h->acc.u32();
    ta.fin = reinterpret_cast<f128*>(h->acc.u32() + h->fin_off());
    timers[n].begin(s, "outer_tail", 0.0);
    hipLaunchKernelGGL(bitz::k_outer_tail, dim3(1), dim3(bitz::kOuterStageBlock), bitz::outer_stage_smem(n - ps.k0), s, ta);
    timers[n].end(s);
    TRY(check_launch());
    h->round = n;
    HIP_TRY(hipMemcpyAsync(h->host.p, h->acc.p, h->acc_words() * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (auto& tmr : timers) tmr.collect();
    return BITZ_OK;
}

// Turns the pipelined eq-factored sums into round coefficients (without
// c1).  With check_claim (round 0 without q(1)) the chain of q(1)
// recoveries started from the caller's claim, and the last round's
// q_{n-1}(r) must equal A(r) B(r) + C(r); with every r_k nonzero a wrong
// claim changes it.  Returns false when the sums cannot be used as they
// are (a recovery would divide by tau_k = 0, or the check fails).
static bool outer_assemble_eq(bitz_outer* h, const bitz_f128* challenges, f128 initial_claim,
                              uint64_t q1mask, bool check_claim, std::vector<bitz_f128>& co) {
    // Grid-wide rounds before kc (a batch member's before k0: no chunk stage).
    const int n = h->n, kc = h->batch_member ? pipe_schedule(n, h->tail_log).k0 : pipe_schedule(n).kc;
    const uint32_t* w = h->host_words();
    for (int k = 0; k < n; ++k)
        if (!((q1mask >> k) & 1) && hostf::is_zero(h->tau[k])) return false;
    h->S = hostf::one();
    h->sigma = initial_claim;
    co.resize(3 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        bitz_f128 raw[3];
        memcpy(raw, w + 12 * k, 48);
        const f128& sc = k >= kc ? h->scale_tail[k] : h->scale_big[k];
        outer_eq_coeffs(h, k, raw, (q1mask >> k) & 1, sc, &co[3 * (size_t)k]);
        outer_eq_advance(h, k, F(challenges[k]));
    }
    if (check_claim) {
        const uint32_t* fin = w + h->fin_off();
        const f128 v = hostf::add(hostf::mul(f128_at(fin), f128_at(fin + 4)), f128_at(fin + 8));
        if (!hostf::eq(v, h->sigma)) return false;
    }
    return true;
}

// The general-factor kernels, all rounds queued back to back; the sums are
// read with the final evaluations.
static int32_t outer_pipeline_generic(bitz_outer* h, const bitz_f128* challenges, std::vector<Timed>& timers,
                                      std::vector<bitz_f128>& co, bitz_f128* claims) {
    hipStream_t s = h->so.s;
    const int n = h->n;
    timers[0].begin(s, "outer_round0", 48.0 * (double)h->len);
    const size_t pairs = h->len / 2;
    hipLaunchKernelGGL(bitz::k_outer_round0, dim3(grid_for(pairs)), dim3(kBlock), 0, s, h->cur[0], h->cur[1],
                       h->cur[2], pairs, h->eqL[h->curL].f(), log2_exact(h->low_len), h->eqH[h->curH].f(),
                       h->acc.u32());
    timers[0].end(s);
    TRY(check_launch());
    for (int k = 0; k < n; ++k) {
        const f128 r = F(challenges[k]);
        const size_t next_len = h->len / 2;
        Dev& dst = (h->where == 1) ? h->Q : h->P;
        const size_t cap = (h->where == 1) ? h->capQ : h->capP;
        f128* out[3] = {dst.f(), dst.f() + cap, dst.f() + 2 * cap};
        uint32_t* slot = h->acc.u32() + 12 * (k + 1);
        if (h->low_len > 1) {
            const size_t nl = h->low_len / 2;
            hipLaunchKernelGGL(bitz::k_fold, dim3(grid_for(nl)), dim3(kBlock), 0, s, h->eqL[h->curL].f(),
                               h->eqL[1 - h->curL].f(), nl, r);
            h->curL ^= 1;
            h->low_len = nl;
            if (next_len > 1) {
                timers[k + 1].begin(s, "outer_fold_next", 72.0 * (double)h->len);
                hipLaunchKernelGGL(bitz::k_outer_fold_next, dim3(grid_for(next_len / 2)), dim3(kBlock), 0, s,
                                   h->cur[0], h->cur[1], h->cur[2], out[0], out[1], out[2], next_len / 2, r,
                                   h->eqL[h->curL].f(), log2_exact(h->low_len), h->eqH[h->curH].f(), slot);
                timers[k + 1].end(s);
            } else {
                hipLaunchKernelGGL(bitz::k_fold3, dim3(1), dim3(kBlock), 0, s, h->cur[0], h->cur[1], h->cur[2],
                                   out[0], out[1], out[2], (size_t)1, r);
            }
        } else {
            const size_t nh = h->high_len / 2;
            hipLaunchKernelGGL(bitz::k_fold, dim3(grid_for(nh)), dim3(kBlock), 0, s, h->eqH[h->curH].f(),
                               h->eqH[1 - h->curH].f(), nh, r);
            h->curH ^= 1;
            h->high_len = nh;
            if (nh == 1) {
                hipLaunchKernelGGL(bitz::k_fold3, dim3(1), dim3(kBlock), 0, s, h->cur[0], h->cur[1], h->cur[2],
                                   out[0], out[1], out[2], (size_t)1, r);
            } else {
                timers[k + 1].begin(s, "outer_fold_next", 72.0 * (double)h->len);
                hipLaunchKernelGGL(bitz::k_outer_fold_next, dim3(grid_for(nh / 2)), dim3(kBlock), 0, s,
                                   h->cur[0], h->cur[1], h->cur[2], out[0], out[1], out[2], nh / 2, r,
                                   h->eqL[h->curL].f(), 0, h->eqH[h->curH].f(), slot);
                timers[k + 1].end(s);
            }
        }
        TRY(check_launch());
        for (int t = 0; t < 3; ++t) h->cur[t] = out[t];
        h->where = (h->where == 1) ? 2 : 1;
        h->len = next_len;
        h->round += 1;
    }
    HIP_TRY(hipMemcpyAsync(h->host.p, h->acc.p, 48 * (size_t)n, hipMemcpyDeviceToHost, s));
    TRY(bitz_outer_finish(h, claims));  // synchronises
    for (auto& tmr : timers) tmr.collect();
    co.assign(static_cast<const bitz_f128*>(h->host.p), static_cast<const bitz_f128*>(h->host.p) + 3 * (size_t)n);
    return BITZ_OK;
}

// prove_outer_sumcheck against challenges fixed in advance (a transcript
// whose challenges do not depend on the messages, such as the seeded
// stream): every round is enqueued back to back and the round sums are read
// once at the end.  Outputs as bitz_prove_outer_sumcheck; eval_points are the
// given challenges.
//
// Eq-factored factors take a first, optimistic pass: round 0 skips q(1)
// (recovered from the claim) unless a challenge is zero, and no later round
// computes it.  If the factor check, a tau_k = 0 or the claim check rejects
// that pass, a second pass computes q(1) wherever the sums need it; factors
// that are not eq tables fall back to the general kernels.
extern "C" int32_t bitz_prove_outer_sumcheck_with_challenges(
    const bitz_f128* challenges, bitz_f128 initial_claim, const bitz_f128* eq_low, uint32_t low_vars,
    const bitz_f128* eq_high, uint32_t high_vars, const bitz_f128* az, uint32_t az_vars, const bitz_f128* bz,
    uint32_t bz_vars, const bitz_f128* cz, uint32_t cz_vars, uint32_t flags, bitz_f128* round_polys,
    bitz_f128* eval_points, bitz_f128* final_claim, bitz_f128* claims) {
    if (!final_claim || !claims || (az_vars && (!round_polys || !eval_points || !challenges)))
        return fail(BITZ_ERR_ARG, "null pointer");
    TRY(outer_check(low_vars, high_vars, az_vars, bz_vars, cz_vars));
    bitz_outer* h = nullptr;
    TRY(outer_begin_impl(&h, eq_low, low_vars, eq_high, high_vars, az, az_vars, bz, bz_vars, cz, cz_vars, flags,
                         false));
    const int n = (int)az_vars;
    int32_t rc = BITZ_OK;
    std::vector<bitz_f128> co;
    f128 claim = F(initial_claim);
    if (n == 0) {
        rc = bitz_outer_finish(h, claims);
        *final_claim = initial_claim;
        bitz_outer_free(h);
        return rc;
    }
    h->started = true;
    if (h->eqmode) {
        std::vector<Timed> timers(n + 1);
        if ((rc = outer_pipeline_prologue(h, challenges))) goto done;
        bool zero_r = false;
        for (int k = 0; k < n; ++k) zero_r |= hostf::is_zero(F(challenges[k]));
        uint64_t q1mask = zero_r ? 1 : 0;
        if ((rc = outer_pipeline_eq(h, q1mask, timers))) goto done;
        if (!outer_validate_eq(h, h->host_words() + h->eq_off())) {
            h->eqmode = false;
            if ((rc = h->rewind()) || (rc = outer_own_factors(h))) goto done;
        } else if (!outer_assemble_eq(h, challenges, claim, q1mask, !(q1mask & 1), co)) {
            q1mask = 1;
            for (int k = 0; k < n; ++k)
                if (hostf::is_zero(h->tau[k])) q1mask |= (uint64_t)1 << k;
            std::vector<Timed> timers2(n + 1);
            if ((rc = h->rewind())) goto done;
            if ((rc = outer_pipeline_eq(h, q1mask, timers2))) goto done;
            if (!outer_assemble_eq(h, challenges, claim, q1mask, false, co)) {
                rc = fail(BITZ_ERR_STATE, "eq-factored sums rejected twice");
                goto done;
            }
        }
        if (h->eqmode) {
            memcpy(claims, h->host_words() + h->fin_off(), 48);
            h->finished = true;
        }
    }
    if (!h->eqmode) {
        std::vector<Timed> timers(n + 1);
        if ((rc = outer_pipeline_generic(h, challenges, timers, co, claims))) goto done;
    }
    for (int k = 0; k < n; ++k) {
        // sumcheck.rs:760-794: c1 = claim + c2 + c3, claim <- g(r).
        f128 c[4] = {F(co[3 * k]), hostf::add(claim, hostf::add(F(co[3 * k + 1]), F(co[3 * k + 2]))),
                     F(co[3 * k + 1]), F(co[3 * k + 2])};
        const f128 r = F(challenges[k]);
        claim = hostf::horner(c, 4, r);
        for (int i = 0; i < 4; ++i) round_polys[4 * k + i] = B(c[i]);
        eval_points[k] = challenges[k];
    }
    *final_claim = B(claim);
done:
    bitz_outer_free(h);
    return rc;
}

// ============================================ batched outer sumchecks ===
// Many independent eq-factored outer sumchecks with challenges fixed in
// advance, proved together: one prologue launch for every instance (factor
// checks, challenge and H(i) tables), one launch per grid-wide round over
// every instance still in its grid-wide rounds (k_batch_r0, k_batch_fold),
// one launch of one-workgroup tails (k_outer_batch), then one read-back.
// The launches are captured in a graph and replayed while the shapes repeat.
// Each instance's result is exactly bitz_prove_outer_sumcheck_with_challenges';
// an instance whose optimistic pass is rejected (factors that are not eq
// tables, tau_k = 0, a wrong claim) is proved again on its own.

namespace {

// A batch member: the eq-factored state of one instance (geometry, tau and
// constants after validation, coefficient recovery) without scratch of its
// own.  Grid-wide rounds run cfg.thread_bits-bit thread grids and end at
// n - cfg.tail_log.
struct BatchConfig {
    int thread_bits, tail_log;
};
int32_t batch_member(bitz_outer** out, const bitz_outer_instance& in, const BatchConfig& cfg) {
    auto* h = new bitz_outer;
    h->batch_member = true;
    h->tail_log = cfg.tail_log;
    h->n = (int)in.num_vars;
    h->N = h->len = (size_t)1 << in.num_vars;
    h->low_len = (size_t)1 << in.low_vars;
    h->high_len = (size_t)1 << in.high_vars;
    h->inputs[0] = reinterpret_cast<const f128*>(in.az);
    h->inputs[1] = reinterpret_cast<const f128*>(in.bz);
    h->inputs[2] = reinterpret_cast<const f128*>(in.cz);
    h->eq_in[0] = reinterpret_cast<const f128*>(in.eq_low);
    h->eq_in[1] = reinterpret_cast<const f128*>(in.eq_high);
    h->eq_bl = eq_check_blocks(h->low_len, 1024);
    h->eq_bh = eq_check_blocks(h->high_len, 1024);
    h->eqmode = outer_setup_eq(h, (int)in.low_vars, (int)in.high_vars, true, cfg.thread_bits);
    *out = h;
    return BITZ_OK;
}

// Tile shapes the autotuner tries for a batch: the thread-grid cap of the
// grid-wide rounds (larger rounds iterate, weighing each iteration through
// its H(i) table) and the size of the one-workgroup tail.
constexpr BatchConfig kBatchDefault = {bitz::kThreadBits, kTailLog};
constexpr BatchConfig kBatchCandidates[] = {{18, 10}, {17, 10}, {16, 10}, {15, 10}, {18, 9}, {17, 9}, {16, 9}};

// Everything a batch of a given shape needs, built once and reused while
// the shapes repeat: member handles (geometry), buffer layout, device
// buffers, pinned staging, the block maps, and the captured launch graph.
// A call refills the descriptors (pointers, q(1) flags) and the challenges
// in the pinned staging; the graph's first node copies them to the device.
struct BatchPlan {
    std::vector<uint32_t> key;  // per instance: num_vars, low_vars, high_vars
    BatchConfig cfg{};
    std::vector<std::unique_ptr<bitz_outer, void (*)(bitz_outer*)>> hs;
    std::vector<uint32_t> members, alone;
    std::vector<int> k0;
    std::vector<size_t> acc_off, ch_off, ht_off, p_off, q_off;
    size_t acc_words = 0, ch_total = 0, ht_total = 0, p_total = 0, q_total = 0;
    int kmax = 0;
    Dev acc, chal, rt4, ht4, pbuf, qbuf, desc;
    void* pin = nullptr;
    size_t pin_bytes = 0;
    // staging layout: [maps][descriptors][challenges][readback]
    size_t pro_m = 0, tl_m = 0, pro_o = 0, tl_o = 0, ch_o = 0, rb_o = 0;
    size_t maps_end = 0, up_end = 0;
    std::vector<size_t> rd_o, rd_m;
    unsigned pro_grid = 0, tail_grid = 0;
    std::vector<unsigned> round_grid;
    std::vector<size_t> round_count;
    StreamOwner so;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchPlan() {
        if (so.s) (void)hipStreamSynchronize(so.s);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pin) (void)hipHostFree(pin);
    }
    uint32_t* host_acc() const { return reinterpret_cast<uint32_t*>(static_cast<char*>(pin) + rb_o); }
};

struct BatchCache {
    std::mutex mu;  // one batch at a time
    std::unique_ptr<BatchPlan> plan;
    std::map<std::vector<uint32_t>, BatchConfig> tuned;
    BatchConfig last{0, 0};
    BatchConfig forced{0, 0};  // set by bitz_outer_batch_set_config
};
// Never destroyed: a cached plan's teardown makes HIP calls, which must not
// run from static destructors after the runtime (or a tool wrapped around it)
// has shut down.  bitz_outer_batch_release frees it explicitly.
BatchCache& g_batch = *new BatchCache;

std::vector<uint32_t> batch_key(const bitz_outer_instance* inst, uint32_t count) {
    std::vector<uint32_t> key;
    key.reserve(3 * (size_t)count);
    for (uint32_t i = 0; i < count; ++i) {
        key.push_back(inst[i].num_vars);
        key.push_back(inst[i].low_vars);
        key.push_back(inst[i].high_vars);
    }
    return key;
}

// Round k's tables of member i: the inputs for k = 0, then the output of
// round k's fold by r_{k-1} (odd rounds in P, even rounds in Q).
const f128* batch_tables(const BatchPlan& P, uint32_t i, int k, int t) {
    if (k == 0) return P.hs[i]->inputs[t];
    if (k & 1) return P.pbuf.f() + t * P.p_total + P.p_off[i];
    return P.qbuf.f() + t * P.q_total + P.q_off[i];
}

int32_t batch_plan_build(BatchPlan& P, const bitz_outer_instance* inst, uint32_t count, const BatchConfig& cfg) {
    P.key = batch_key(inst, count);
    P.cfg = cfg;
    for (uint32_t i = 0; i < count; ++i) {
        bitz_outer* h = nullptr;
        TRY(batch_member(&h, inst[i], cfg));
        P.hs.emplace_back(h, [](bitz_outer* p) { delete p; });
        if (h->eqmode && h->n >= 1 && h->n <= bitz::kMaxRounds)
            P.members.push_back(i);
        else
            P.alone.push_back(i);
    }
    if (P.members.empty()) return BITZ_OK;
    TRY(P.so.create());
    P.k0.assign(count, 0);
    P.acc_off.assign(count, 0);
    P.ch_off.assign(count, 0);
    P.ht_off.assign(count, 0);
    P.p_off.assign(count, 0);
    P.q_off.assign(count, 0);
    for (uint32_t i : P.members) {
        bitz_outer* h = P.hs[i].get();
        P.k0[i] = pipe_schedule(h->n, cfg.tail_log).k0;
        P.kmax = std::max(P.kmax, P.k0[i]);
        P.acc_off[i] = P.acc_words;
        P.acc_words += (h->acc_words() + 63) & ~(size_t)63;
        P.ch_off[i] = P.ch_total;
        P.ch_total += (size_t)h->n;
        P.ht_off[i] = P.ht_total;
        P.ht_total += h->htab_total;
        P.p_off[i] = P.p_total;
        if (P.k0[i] >= 2) P.p_total += h->N / 2;
        P.q_off[i] = P.q_total;
        if (P.k0[i] >= 3) P.q_total += h->N / 4;
    }
    TRY(P.acc.alloc(P.acc_words * 4));
    TRY(P.chal.alloc(std::max<size_t>(P.ch_total, 1) * sizeof(f128)));
    TRY(P.rt4.alloc(std::max<size_t>(P.ch_total, 1) * kConstTable4Entries * sizeof(uint4)));
    TRY(P.ht4.alloc(std::max<size_t>(P.ht_total, 1) * kConstTable4Entries * sizeof(uint4)));
    TRY(P.pbuf.alloc(std::max<size_t>(3 * P.p_total, 1) * sizeof(f128)));
    TRY(P.qbuf.alloc(std::max<size_t>(3 * P.q_total, 1) * sizeof(f128)));
    // Block maps: workgroup -> descriptor of every launch.
    std::vector<uint32_t> pro_map, tail_map;
    std::vector<std::vector<uint32_t>> round_map(P.kmax);
    P.round_count.assign(P.kmax, 0);
    for (uint32_t j = 0; j < (uint32_t)P.members.size(); ++j) {
        bitz_outer* h = P.hs[P.members[j]].get();
        pro_map.insert(pro_map.end(), h->eq_bl + h->eq_bh + (unsigned)h->n + h->htab_total, j);
        tail_map.push_back(j);
    }
    for (int k = 0; k < P.kmax; ++k) {
        const unsigned block = k == 0 ? bitz::kR0Block : bitz::kFoldBlock;
        for (uint32_t i : P.members) {
            if (k >= P.k0[i]) continue;
            const auto& g = P.hs[i]->geo[k];
            round_map[k].insert(round_map[k].end(), (g.threads + block - 1) / block, (uint32_t)P.round_count[k]);
            P.round_count[k] += 1;
        }
    }
    size_t off = 0;
    auto place = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        const size_t o = off;
        off += bytes;
        return o;
    };
    P.pro_m = place(pro_map.size() * 4);
    P.rd_m.resize(P.kmax);
    for (int k = 0; k < P.kmax; ++k) P.rd_m[k] = place(round_map[k].size() * 4);
    P.tl_m = place(tail_map.size() * 4);
    P.maps_end = (off + 255) & ~(size_t)255;
    P.pro_o = place(P.members.size() * sizeof(bitz::BatchPrologue));
    P.rd_o.resize(P.kmax);
    for (int k = 0; k < P.kmax; ++k) P.rd_o[k] = place(P.round_count[k] * sizeof(bitz::RoundDesc));
    P.tl_o = place(P.members.size() * sizeof(bitz::LdsStage));
    P.ch_o = place(P.ch_total * sizeof(f128));
    P.up_end = off;
    P.rb_o = place(P.acc_words * 4);
    P.pin_bytes = off;
    if (hipHostMalloc(&P.pin, P.pin_bytes, hipHostMallocDefault) != hipSuccess) {
        P.pin = nullptr;
        return fail(BITZ_ERR_ALLOC, "hipHostMalloc failed");
    }
    char* pin = static_cast<char*>(P.pin);
    memcpy(pin + P.pro_m, pro_map.data(), pro_map.size() * 4);
    for (int k = 0; k < P.kmax; ++k) memcpy(pin + P.rd_m[k], round_map[k].data(), round_map[k].size() * 4);
    memcpy(pin + P.tl_m, tail_map.data(), tail_map.size() * 4);
    TRY(P.desc.alloc(P.up_end));
    HIP_TRY(hipMemcpyAsync(P.desc.p, pin, P.maps_end, hipMemcpyHostToDevice, P.so.s));
    HIP_TRY(hipStreamSynchronize(P.so.s));
    P.pro_grid = (unsigned)pro_map.size();
    P.tail_grid = (unsigned)tail_map.size();
    P.round_grid.resize(P.kmax);
    for (int k = 0; k < P.kmax; ++k) P.round_grid[k] = (unsigned)round_map[k].size();
    return BITZ_OK;
}

// Writes this call's descriptors and challenges into the plan's staging.
void batch_plan_fill(BatchPlan& P, const bitz_outer_instance* inst, const std::vector<uint64_t>& q1mask) {
    char* pin = static_cast<char*>(P.pin);
    auto rtab_of = [&](uint32_t i) { return static_cast<uint4*>(P.rt4.p) + P.ch_off[i] * kConstTable4Entries; };
    for (uint32_t i : P.members) {
        bitz_outer* h = P.hs[i].get();
        const auto& x = inst[i];
        h->inputs[0] = reinterpret_cast<const f128*>(x.az);
        h->inputs[1] = reinterpret_cast<const f128*>(x.bz);
        h->inputs[2] = reinterpret_cast<const f128*>(x.cz);
        h->eq_in[0] = reinterpret_cast<const f128*>(x.eq_low);
        h->eq_in[1] = reinterpret_cast<const f128*>(x.eq_high);
        memcpy(pin + P.ch_o + P.ch_off[i] * sizeof(f128), x.challenges, x.num_vars * sizeof(f128));
    }
    auto* pro = reinterpret_cast<bitz::BatchPrologue*>(pin + P.pro_o);
    auto* tails = reinterpret_cast<bitz::LdsStage*>(pin + P.tl_o);
    uint32_t pro_base = 0;
    for (uint32_t j = 0; j < (uint32_t)P.members.size(); ++j) {
        const uint32_t i = P.members[j];
        bitz_outer* h = P.hs[i].get();
        bitz::BatchPrologue p{};
        p.L = h->eq_in[0];
        p.H = h->eq_in[1];
        p.ell = h->ell;
        p.m = h->m;
        p.n = h->n;
        p.zero = P.acc.u32() + P.acc_off[i];
        p.zero_words = (unsigned)h->eq_off();
        p.eq_blocks_l = h->eq_bl;
        p.eq_blocks = h->eq_bl + h->eq_bh;
        p.eq_out = P.acc.u32() + P.acc_off[i] + h->eq_off();
        p.chal = P.chal.f() + P.ch_off[i];
        p.rtabs4 = rtab_of(i);
        p.hdesc = h->hdesc;
        p.htabs4 = static_cast<uint4*>(P.ht4.p) + P.ht_off[i] * kConstTable4Entries;
        p.block_base = pro_base;
        pro_base += p.eq_blocks + (unsigned)h->n + (unsigned)h->htab_total;
        pro[j] = p;
        // Tail: round k0-1's tables folded by r_{k0-1} on load (the inputs when k0 = 0).
        const int kk = P.k0[i];
        const int src = kk > 0 ? kk - 1 : 0;
        bitz::LdsStage t{};
        t.a = batch_tables(P, i, src, 0);
        t.b = batch_tables(P, i, src, 1);
        t.c = batch_tables(P, i, src, 2);
        t.L = h->eq_in[0];
        t.H = h->eq_in[1];
        t.rtabs4 = rtab_of(i);
        t.acc = P.acc.u32() + P.acc_off[i];
        t.fin = reinterpret_cast<f128*>(P.acc.u32() + P.acc_off[i] + h->fin_off());
        t.q1mask = q1mask[i];
        t.ell = h->ell;
        t.n = h->n;
        t.k_first = kk;
        t.s = h->n - kk;
        t.lg_in = kk > 0 ? h->n - kk + 1 : h->n;
        t.fold_first = kk > 0;
        t.final_stage = 1;
        t.store_sums = 1;
        t.chunk_base = j;  // one workgroup per member: chunk 0
        tails[j] = t;
    }
    for (int k = 0; k < P.kmax; ++k) {
        auto* rd = reinterpret_cast<bitz::RoundDesc*>(pin + P.rd_o[k]);
        const unsigned block = k == 0 ? bitz::kR0Block : bitz::kFoldBlock;
        uint32_t base = 0, c = 0;
        for (uint32_t i : P.members) {
            if (k >= P.k0[i]) continue;
            bitz_outer* h = P.hs[i].get();
            const auto& g = h->geo[k];
            bitz::RoundDesc d{};
            // Round k reads round k-1's tables and folds them by r_{k-1}.
            const int src = k == 0 ? 0 : k - 1;
            d.a = batch_tables(P, i, src, 0);
            d.b = batch_tables(P, i, src, 1);
            d.c = batch_tables(P, i, src, 2);
            if (k > 0) {
                d.ao = const_cast<f128*>(batch_tables(P, i, k, 0));
                d.bo = const_cast<f128*>(batch_tables(P, i, k, 1));
                d.co = const_cast<f128*>(batch_tables(P, i, k, 2));
                d.r_tab4 = rtab_of(i) + (size_t)(k - 1) * kConstTable4Entries;
            }
            d.la = h->table_ptr(g.la.table);
            d.lb = h->table_ptr(g.lb.table);
            d.la_bits = g.la_bits;
            d.la_shift = g.la.a - h->table_base(g.la.table);
            d.lb_shift = g.lb.a - h->table_base(g.lb.table);
            d.h_tabs4 = static_cast<const uint4*>(P.ht4.p) + (P.ht_off[i] + g.htab_off) * kConstTable4Entries;
            d.acc = P.acc.u32() + P.acc_off[i] + 12 * (size_t)k;
            d.threads = g.threads;
            d.iters = g.iters;
            d.q1 = (int)((q1mask[i] >> k) & 1);
            d.block_base = base;
            base += (uint32_t)((g.threads + block - 1) / block);
            rd[c++] = d;
        }
    }
}

// Enqueues one batch on the plan's stream: descriptors and challenges up,
// prologue, grid-wide rounds, tails, sums down.  Per-kernel timers only when
// not captured.
int32_t batch_plan_enqueue(BatchPlan& P, bool timed) {
    hipStream_t s = P.so.s;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_outer_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::stage_smem(bitz::kStageMaxLog));
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_batch_r0),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kR0Smem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_batch_fold),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kFoldSmem);
        attr = true;
    }
    char* pin = static_cast<char*>(P.pin);
    const char* dp = static_cast<const char*>(P.desc.p);
    HIP_TRY(hipMemcpyAsync(static_cast<char*>(P.desc.p) + P.maps_end, pin + P.maps_end, P.ch_o - P.maps_end,
                           hipMemcpyHostToDevice, s));
    if (P.ch_total)
        HIP_TRY(hipMemcpyAsync(P.chal.p, pin + P.ch_o, P.ch_total * sizeof(f128), hipMemcpyHostToDevice, s));
    std::vector<Timed> ts(timed ? P.kmax + 2 : 0);
    if (timed) ts[0].begin(s, "outer_batch_prologue", 0.0);
    hipLaunchKernelGGL(bitz::k_batch_prologue, dim3(P.pro_grid), dim3(1024), 0, s,
                       reinterpret_cast<const bitz::BatchPrologue*>(dp + P.pro_o),
                       reinterpret_cast<const uint32_t*>(dp + P.pro_m));
    if (timed) ts[0].end(s);
    TRY(check_launch());
    for (int k = 0; k < P.kmax; ++k) {
        const auto* d = reinterpret_cast<const bitz::RoundDesc*>(dp + P.rd_o[k]);
        const auto* m = reinterpret_cast<const uint32_t*>(dp + P.rd_m[k]);
        if (timed) ts[k + 1].begin(s, k == 0 ? "outer_batch_round0" : "outer_batch_fold", 0.0);
        if (k == 0)
            hipLaunchKernelGGL(bitz::k_batch_r0, dim3(P.round_grid[k]), dim3(bitz::kR0Block), bitz::kR0Smem, s, d, m);
        else
            hipLaunchKernelGGL(bitz::k_batch_fold, dim3(P.round_grid[k]), dim3(bitz::kFoldBlock), bitz::kFoldSmem, s,
                               d, m);
        if (timed) ts[k + 1].end(s);
        TRY(check_launch());
    }
    if (timed) ts[P.kmax + 1].begin(s, "outer_batch_tail", 0.0);
    hipLaunchKernelGGL(bitz::k_outer_batch, dim3(P.tail_grid), dim3(bitz::kStageBlock),
                       bitz::stage_smem(bitz::kStageMaxLog), s, reinterpret_cast<const bitz::LdsStage*>(dp + P.tl_o),
                       reinterpret_cast<const uint32_t*>(dp + P.tl_m));
    if (timed) ts[P.kmax + 1].end(s);
    TRY(check_launch());
    HIP_TRY(hipMemcpyAsync(P.host_acc(), P.acc.p, P.acc_words * 4, hipMemcpyDeviceToHost, s));
    if (timed) {
        HIP_TRY(hipStreamSynchronize(s));
        for (auto& t : ts) t.collect();
    }
    return BITZ_OK;
}

// Runs the plan once: through its captured graph (captured on first use)
// unless eager, then waits for the sums.
int32_t batch_plan_run(BatchPlan& P, bool eager) {
    hipStream_t s = P.so.s;
    if (eager || g_prof.on) {
        TRY(batch_plan_enqueue(P, true));
        HIP_TRY(hipStreamSynchronize(s));
        return BITZ_OK;
    }
    if (!P.exec) {
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int32_t rc = batch_plan_enqueue(P, false);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return fail(BITZ_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        P.graph = g;
        HIP_TRY(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(P.exec, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BITZ_OK;
}

// q(1) in round 0 when a challenge is zero (otherwise the claim check at
// the end catches a wrong claim).
std::vector<uint64_t> batch_q1mask(const BatchPlan& P, const bitz_outer_instance* inst, uint32_t count) {
    std::vector<uint64_t> q1mask(count, 0);
    for (uint32_t i : P.members)
        for (uint32_t k = 0; k < inst[i].num_vars; ++k)
            if (hostf::is_zero(F(inst[i].challenges[k]))) q1mask[i] = 1;
    return q1mask;
}

// Times each candidate shape on this batch (one warm run, then one timed
// run each) and returns the fastest.
int32_t batch_autotune(const bitz_outer_instance* inst, uint32_t count, bool eager, BatchConfig* best) {
    float best_ms = 0;
    *best = kBatchDefault;
    bool any = false;
    for (const BatchConfig& c : kBatchCandidates) {
        BatchPlan P;
        TRY(batch_plan_build(P, inst, count, c));
        if (P.members.empty()) return BITZ_OK;
        batch_plan_fill(P, inst, batch_q1mask(P, inst, count));
        TRY(batch_plan_run(P, eager));
        const auto t0 = std::chrono::steady_clock::now();
        TRY(batch_plan_run(P, eager));
        const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!any || ms < best_ms) {
            best_ms = ms;
            *best = c;
            any = true;
        }
    }
    return BITZ_OK;
}

}  // namespace

extern "C" int32_t bitz_prove_outer_batch(const bitz_outer_instance* inst, uint32_t count, uint32_t flags) {
    const auto t_total = std::chrono::steady_clock::now();
    struct TotalTime {  // host phases of a batch call, with profiling on
        std::chrono::steady_clock::time_point t0;
        ~TotalTime() { prof_add_host("batch_total", t0); }
    } total_time{t_total};
    if (!inst && count) return fail(BITZ_ERR_ARG, "null pointer");
    if (!(flags & BITZ_DEVICE)) return fail(BITZ_ERR_ARG, "batched proofs take device tables (BITZ_DEVICE)");
    for (uint32_t i = 0; i < count; ++i) {
        const auto& x = inst[i];
        if (!x.final_claim || !x.claims || (x.num_vars && (!x.round_polys || !x.eval_points || !x.challenges)) ||
            !x.eq_low || !x.eq_high || !x.az || !x.bz || !x.cz)
            return fail(BITZ_ERR_ARG, "instance " + std::to_string(i) + ": null pointer");
        const int32_t rc = outer_check(x.low_vars, x.high_vars, x.num_vars, x.num_vars, x.num_vars);
        if (rc) return fail(rc, "instance " + std::to_string(i) + ": " + g_err);
    }
    if (!count) return BITZ_OK;
    std::lock_guard<std::mutex> lock(g_batch.mu);
    const bool eager = (flags & BITZ_BATCH_EAGER) != 0;
    const std::vector<uint32_t> key = batch_key(inst, count);
    BatchConfig cfg = kBatchDefault;
    if (g_batch.forced.thread_bits) {
        cfg = g_batch.forced;
    } else if (!(flags & BITZ_BATCH_NO_TUNE)) {
        auto it = g_batch.tuned.find(key);
        if (it == g_batch.tuned.end()) {
            g_batch.plan.reset();  // its buffers return to the cache for the trials
            TRY(batch_autotune(inst, count, eager, &cfg));
            it = g_batch.tuned.emplace(key, cfg).first;
        }
        cfg = it->second;
    }
    g_batch.last = cfg;
    if (!g_batch.plan || g_batch.plan->key != key || g_batch.plan->cfg.thread_bits != cfg.thread_bits ||
        g_batch.plan->cfg.tail_log != cfg.tail_log) {
        g_batch.plan.reset();
        auto P = std::make_unique<BatchPlan>();
        TRY(batch_plan_build(*P, inst, count, cfg));
        g_batch.plan = std::move(P);
    }
    BatchPlan& P = *g_batch.plan;
    std::vector<uint32_t> alone = P.alone;
    int32_t rc = BITZ_OK;
    if (!P.members.empty()) {
        const auto t_fill = std::chrono::steady_clock::now();
        const std::vector<uint64_t> q1mask = batch_q1mask(P, inst, count);
        batch_plan_fill(P, inst, q1mask);
        prof_add_host("batch_fill", t_fill);
        const auto t_run = std::chrono::steady_clock::now();
        TRY(batch_plan_run(P, eager));
        prof_add_host("batch_run", t_run);
        const auto t_finish = std::chrono::steady_clock::now();
        struct FinishTime {
            std::chrono::steady_clock::time_point t0;
            ~FinishTime() { prof_add_host("batch_finish", t0); }
        } finish_time{t_finish};
        const uint32_t* host_acc = P.host_acc();
        // Per instance: validate, recover the coefficients, check the claim.
        // Instances are independent; a few host threads share them.
        auto finish = [&](uint32_t i) -> bool {
            bitz_outer* h = P.hs[i].get();
            h->host_ext = host_acc + P.acc_off[i];
            const auto& x = inst[i];
            std::vector<bitz_f128> co;
            const bool ok = outer_validate_eq(h, h->host_words() + h->eq_off()) &&
                            outer_assemble_eq(h, x.challenges, F(x.initial_claim), q1mask[i], !(q1mask[i] & 1), co);
            if (!ok) return false;
            f128 claim = F(x.initial_claim);
            for (int k = 0; k < h->n; ++k) {
                f128 c[4] = {F(co[3 * k]), hostf::add(claim, hostf::add(F(co[3 * k + 1]), F(co[3 * k + 2]))),
                             F(co[3 * k + 1]), F(co[3 * k + 2])};
                const f128 r = F(x.challenges[k]);
                claim = hostf::horner(c, 4, r);
                for (int t = 0; t < 4; ++t) x.round_polys[4 * k + t] = B(c[t]);
                x.eval_points[k] = x.challenges[k];
            }
            *x.final_claim = B(claim);
            memcpy(x.claims, h->host_words() + h->fin_off(), 48);
            return true;
        };
        const auto& members = P.members;
        std::vector<char> done(members.size(), 0);
        const size_t workers = std::min<size_t>({8, std::max(1u, std::thread::hardware_concurrency()),
                                                 (members.size() + 31) / 32});
        std::vector<std::thread> pool;
        for (size_t w = 1; w < workers; ++w)
            pool.emplace_back([&, w] {
                for (size_t j = w; j < members.size(); j += workers) done[j] = finish(members[j]);
            });
        for (size_t j = 0; j < members.size(); j += workers) done[j] = finish(members[j]);
        for (auto& t : pool) t.join();
        for (size_t j = 0; j < members.size(); ++j)
            if (!done[j]) alone.push_back(members[j]);
    }
    for (uint32_t i : alone) {
        const auto& x = inst[i];
        rc = bitz_prove_outer_sumcheck_with_challenges(x.challenges, x.initial_claim, x.eq_low, x.low_vars,
                                                       x.eq_high, x.high_vars, x.az, x.num_vars, x.bz, x.num_vars,
                                                       x.cz, x.num_vars, flags & BITZ_DEVICE, x.round_polys,
                                                       x.eval_points, x.final_claim, x.claims);
        if (rc) return fail(rc, "instance " + std::to_string(i) + ": " + g_err);
    }
    return rc;
}

extern "C" int32_t bitz_outer_batch_config(uint32_t* thread_bits, uint32_t* tail_log) {
    std::lock_guard<std::mutex> lock(g_batch.mu);
    if (thread_bits) *thread_bits = (uint32_t)g_batch.last.thread_bits;
    if (tail_log) *tail_log = (uint32_t)g_batch.last.tail_log;
    return BITZ_OK;
}

extern "C" int32_t bitz_outer_batch_set_config(uint32_t thread_bits, uint32_t tail_log) {
    if (thread_bits && (thread_bits < 6 || thread_bits > 20 || tail_log < 1 || (int)tail_log > kTailLog))
        return fail(BITZ_ERR_ARG, "batch tile shapes out of range");
    std::lock_guard<std::mutex> lock(g_batch.mu);
    g_batch.forced = thread_bits ? BatchConfig{(int)thread_bits, (int)tail_log} : BatchConfig{0, 0};
    return BITZ_OK;
}

extern "C" void bitz_outer_batch_release(void) {
    std::lock_guard<std::mutex> lock(g_batch.mu);
    g_batch.plan.reset();
}

// ======================================================= inner sumcheck ===
struct bitz_inner {
    int n = 0;
    size_t len = 0;
    // Entries of the current tables that may be nonzero: from round 0 on,
    // 1 + the last index where D or h is nonzero, halved (rounded up) by
    // each fold; the fold rounds read nothing past it.
    size_t live = 0;
    int round = 0;
    bool started = false;
    Dev in_d, in_w, P, Q, rtab;  // rtab: the current challenge's 4-bit table
    const f128* cur[2] = {nullptr, nullptr};
    int where = 0;
    size_t capP = 0, capQ = 0;
    Dev acc;
    Pinned host;
    Timed tm;
    // Packed mode (bitz_inner_begin_packed): h is a Boolean assignment given
    // as its packed bits.  Rounds 0 .. kb run the bit kernels (k_inner_bits:
    // h_K in the monomial basis of r_0 .. r_{K-1}, sums multiplied by the
    // monomials here), round kb also writes h_kb as elements, and the
    // general kernels take the later rounds.
    const uint64_t* bits = nullptr;
    Dev in_bits, bacc;  // bacc: round K's sums (kInnerBitsAccWords words) at bacc + kInnerBitsAccWords K
    int kb = -1;
    std::vector<f128> r;  // the challenges so far
    StreamOwner so;  // last: destroyed (and synchronised) before the buffers
};

// Grid of the inner sumcheck's lane-table kernels (grid-stride loops).
static unsigned inner_grid(size_t work) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>((work + bitz::kInnerBlock - 1) / bitz::kInnerBlock, 2048));
}
static void inner_set_attributes() {
    static bool attr = false;
    if (attr) return;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_inner_r0_lane),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kInnerGridSmem);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_inner_fold_lane),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kInnerGridSmem);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_inner_tail),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kInnerTailSmem);
    attr = true;
}

// r_S = prod_{i in S} r_i for every S < 2^K (bit i of S selects r_i).
static void monomials(const f128* r, int K, f128* mono) {
    mono[0] = hostf::one();
    for (int i = 0; i < K; ++i)
        for (int S = 0; S < (1 << i); ++S) mono[S | (1 << i)] = hostf::mul(mono[S], r[i]);
}

// Round K's [c0, c2] from a bit kernel's sums (the XOR of their copies): 2^K
// sums for c0, then 2^K for c2, each weighed by its monomial r_S.  *live
// (round 0) receives the largest live length the copies hold.
static void bits_round_coeffs(const uint32_t* acc, int K, const f128* r, bitz_f128* co,
                              unsigned long long* live = nullptr) {
    const int S = 1 << K;
    uint32_t w[8 << bitz::kInnerBitsMaxRound] = {};
    unsigned long long lv = 0;
    for (int c = 0; c < bitz::kInnerBitsReplicas; ++c) {
        const uint32_t* rep = acc + (size_t)c * bitz::kInnerBitsReplicaWords;
        for (int i = 0; i < 8 * S; ++i) w[i] ^= rep[i];
        unsigned long long x;
        memcpy(&x, rep + bitz::kInnerBitsLiveWord, 8);
        lv = std::max(lv, x);
    }
    if (live) *live = lv;
    f128 mono[1 << bitz::kInnerBitsMaxRound];
    monomials(r, K, mono);
    f128 c0{0, 0}, c2{0, 0};
    for (int s = 0; s < S; ++s) {
        c0 = hostf::add(c0, hostf::mul(mono[s], f128_at(w + 4 * s)));
        c2 = hostf::add(c2, hostf::mul(mono[s], f128_at(w + 4 * (S + s))));
    }
    co[0] = B(c0);
    co[1] = B(c2);
}

// Workgroups of a bit round over `pairs` output pairs.
static unsigned bits_grid(size_t pairs) {
    return (unsigned)std::max<size_t>(
        1, std::min<size_t>((pairs + bitz::kInnerBlock - 1) / bitz::kInnerBlock, bitz::kInnerBitsMaxBlocks));
}

// Launches round K's bit kernel (one inner sumcheck, grid-stride).
static int32_t launch_inner_bits(int K, const bitz::InnerBitsDesc& A, unsigned grid, hipStream_t s) {
    static bool attr = false;
    if (!attr) {
        const void* fns[] = {reinterpret_cast<const void*>(&bitz::k_inner_bits<1>),
                             reinterpret_cast<const void*>(&bitz::k_inner_bits<2>),
                             reinterpret_cast<const void*>(&bitz::k_inner_bits<3>),
                             reinterpret_cast<const void*>(&bitz::k_inner_bits_batch<1>),
                             reinterpret_cast<const void*>(&bitz::k_inner_bits_batch<2>),
                             reinterpret_cast<const void*>(&bitz::k_inner_bits_batch<3>)};
        for (const void* f : fns)
            (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kInnerBitsSmem);
        attr = true;
    }
    const size_t smem = K ? bitz::kInnerBitsSmem : 0;
    switch (K) {
        case 0: hipLaunchKernelGGL(bitz::k_inner_bits<0>, dim3(grid), dim3(bitz::kInnerBlock), smem, s, A); break;
        case 1: hipLaunchKernelGGL(bitz::k_inner_bits<1>, dim3(grid), dim3(bitz::kInnerBlock), smem, s, A); break;
        case 2: hipLaunchKernelGGL(bitz::k_inner_bits<2>, dim3(grid), dim3(bitz::kInnerBlock), smem, s, A); break;
        case 3: hipLaunchKernelGGL(bitz::k_inner_bits<3>, dim3(grid), dim3(bitz::kInnerBlock), smem, s, A); break;
        default: return fail(BITZ_ERR_ARG, "no bit kernel for this round");
    }
    return check_launch();
}

// The handle over D and h (witness null: the packed mode sets h up).
static int32_t inner_begin_impl(bitz_inner** out, const bitz_f128* matrix, const bitz_f128* witness,
                                uint32_t matrix_vars, uint32_t flags) {
    auto* h = new bitz_inner;
    int32_t rc;
    auto cleanup = [&](int32_t c) {
        delete h;
        return c;
    };
    if ((rc = h->so.create())) return cleanup(rc);
    h->n = (int)matrix_vars;
    const size_t N = (size_t)1 << matrix_vars;
    h->len = N;
    h->live = N;
    if (flags & BITZ_DEVICE) {
        h->cur[0] = reinterpret_cast<const f128*>(matrix);
        h->cur[1] = reinterpret_cast<const f128*>(witness);
    } else {
        if ((rc = upload(h->in_d, matrix, N, flags, h->so.s))) return cleanup(rc);
        h->cur[0] = h->in_d.f();
        if (witness) {
            if ((rc = upload(h->in_w, witness, N, flags, h->so.s))) return cleanup(rc);
            h->cur[1] = h->in_w.f();
        }
    }
    h->capP = std::max<size_t>(N / 2, 1);
    h->capQ = std::max<size_t>(N / 4, 1);
    if ((rc = h->P.alloc(2 * h->capP * sizeof(f128)))) return cleanup(rc);
    if ((rc = h->Q.alloc(2 * h->capQ * sizeof(f128)))) return cleanup(rc);
    if ((rc = h->acc.alloc((size_t)(h->n + 2) * 8 * sizeof(uint32_t)))) return cleanup(rc);
    if (hipMemsetAsync(h->acc.p, 0, h->acc.bytes, h->so.s) != hipSuccess)
        return cleanup(fail(BITZ_ERR_DEVICE, "memset failed"));
    if ((rc = h->host.alloc(bitz::kInnerBitsAccWords * sizeof(uint32_t)))) return cleanup(rc);
    *out = h;
    return BITZ_OK;
}

extern "C" int32_t bitz_inner_begin(bitz_inner** out, const bitz_f128* matrix, uint32_t matrix_vars,
                                    const bitz_f128* witness, uint32_t witness_vars, uint32_t flags) {
    if (!out || !matrix || !witness) return fail(BITZ_ERR_ARG, "null pointer");
    if (witness_vars != matrix_vars)
        return fail(BITZ_ERR_PRODUCT_DIMS, "matrix and witness have different num_vars");
    if (matrix_vars > 34) return fail(BITZ_ERR_ARG, "num_vars too large");
    return inner_begin_impl(out, matrix, witness, matrix_vars, flags);
}

extern "C" int32_t bitz_inner_begin_packed(bitz_inner** out, const bitz_f128* matrix, uint32_t matrix_vars,
                                           const uint64_t* bits, uint32_t assignment_vars, uint32_t flags) {
    if (!out || !matrix || !bits) return fail(BITZ_ERR_ARG, "null pointer");
    if (assignment_vars != matrix_vars)
        return fail(BITZ_ERR_PRODUCT_DIMS, "matrix and assignment have different num_vars");
    if (matrix_vars > 34) return fail(BITZ_ERR_ARG, "num_vars too large");
    TRY(inner_begin_impl(out, matrix, nullptr, matrix_vars, flags));
    bitz_inner* h = *out;
    auto cleanup = [&](int32_t c) {
        bitz_inner_free(h);
        *out = nullptr;
        return c;
    };
    const size_t N = (size_t)1 << matrix_vars;
    const size_t words = (N + 63) / 64;
    int32_t rc;
    h->bits = bits;
    if (!(flags & BITZ_DEVICE)) {
        if ((rc = h->in_bits.alloc(words * 8))) return cleanup(rc);
        if (hipMemcpyAsync(h->in_bits.p, bits, words * 8, hipMemcpyHostToDevice, h->so.s) != hipSuccess)
            return cleanup(fail(BITZ_ERR_DEVICE, "copy of the assignment bits failed"));
        h->bits = static_cast<const uint64_t*>(h->in_bits.p);
    }
    if (matrix_vars <= 1) {
        // One round at most: the assignment as elements and the general kernels.
        if ((rc = h->in_w.alloc(N * sizeof(f128)))) return cleanup(rc);
        hipLaunchKernelGGL(bitz::k_assignment_bits, dim3(1), dim3(kBlock), 0, h->so.s, h->bits, N, h->in_w.f(), N);
        if ((rc = check_launch())) return cleanup(rc);
        h->cur[1] = h->in_w.f();
        h->bits = nullptr;
        return BITZ_OK;
    }
    h->kb = std::min<int>(bitz::kInnerBitsMaxRound, (int)matrix_vars - 1);
    if ((rc = h->bacc.alloc((size_t)(bitz::kInnerBitsMaxRound + 1) * bitz::kInnerBitsAccWords * sizeof(uint32_t))))
        return cleanup(rc);
    if (hipMemsetAsync(h->bacc.p, 0, h->bacc.bytes, h->so.s) != hipSuccess)
        return cleanup(fail(BITZ_ERR_DEVICE, "memset failed"));
    return BITZ_OK;
}

static int32_t inner_read(bitz_inner* h, int slot, bitz_f128* co) {
    HIP_TRY(hipMemcpyAsync(h->host.p, h->acc.u32() + 8 * slot, 32, hipMemcpyDeviceToHost, h->so.s));
    HIP_TRY(hipStreamSynchronize(h->so.s));
    memcpy(co, h->host.p, 32);
    h->tm.collect();
    return BITZ_OK;
}

extern "C" int32_t bitz_inner_round0(bitz_inner* h, bitz_f128* co) {
    if (!h || !co) return fail(BITZ_ERR_ARG, "null pointer");
    if (h->started) return fail(BITZ_ERR_STATE, "round0 called twice");
    h->started = true;
    if (h->n == 0) {
        memset(co, 0, 32);
        return BITZ_OK;
    }
    // Round 0 reads both tables in full and records their live length in
    // the spare slot n + 1 (slot k holds round k's sums, k < n): the fold
    // rounds then skip the tail where D and h are both zero.
    const size_t pairs = h->len / 2;
    auto* live_word = reinterpret_cast<unsigned long long*>(h->acc.u32() + 8 * (h->n + 1));
    inner_set_attributes();
    if (h->bits) {
        // Round 0 over the bits: D read in full, [c0, c2] by masked XORs.
        bitz::InnerBitsDesc A{};
        A.d = h->cur[0];
        A.bits = h->bits;
        A.pairs = pairs;
        A.live = h->len;
        A.acc = h->bacc.u32();
        A.blocks = bits_grid(pairs);
        h->tm.begin(h->so.s, "inner_round0", 16.0 * (double)h->len + (double)h->len / 8);
        TRY(launch_inner_bits(0, A, A.blocks, h->so.s));
        h->tm.end(h->so.s);
        HIP_TRY(hipMemcpyAsync(h->host.p, h->bacc.p, bitz::kInnerBitsAccWords * 4, hipMemcpyDeviceToHost, h->so.s));
        HIP_TRY(hipStreamSynchronize(h->so.s));
        h->tm.collect();
        unsigned long long live;
        bits_round_coeffs(static_cast<const uint32_t*>(h->host.p), 0, nullptr, co, &live);
        h->live = std::max<size_t>(1, std::min<size_t>((size_t)live, h->len));
        return BITZ_OK;
    }
    h->tm.begin(h->so.s, "inner_round0", 32.0 * (double)h->len);
    hipLaunchKernelGGL(bitz::k_inner_r0_lane, dim3(inner_grid(pairs)), dim3(bitz::kInnerBlock),
                       bitz::kInnerGridSmem, h->so.s, h->cur[0], h->cur[1], pairs, h->acc.u32(), live_word);
    h->tm.end(h->so.s);
    TRY(check_launch());
    HIP_TRY(hipMemcpyAsync(static_cast<char*>(h->host.p) + 32, live_word, 8, hipMemcpyDeviceToHost, h->so.s));
    TRY(inner_read(h, 0, co));
    unsigned long long live;
    memcpy(&live, static_cast<const char*>(h->host.p) + 32, 8);
    h->live = std::max<size_t>(1, std::min<size_t>((size_t)live, h->len));
    return BITZ_OK;
}

extern "C" int32_t bitz_inner_fold_next(bitz_inner* h, bitz_f128 rb, bitz_f128* co) {
    if (!h || !co) return fail(BITZ_ERR_ARG, "null pointer");
    if (!h->started || h->round >= h->n) return fail(BITZ_ERR_STATE, "no round left to fold");
    hipStream_t s = h->so.s;
    const f128 r = F(rb);
    const size_t next_len = h->len / 2;
    Dev& dst = (h->where == 1) ? h->Q : h->P;
    const size_t cap = (h->where == 1) ? h->capQ : h->capP;
    f128* out[2] = {dst.f(), dst.f() + cap};
    bool has_next = false;
    const int K = h->round + 1;  // the round whose sums this call computes
    if (h->bits && K <= h->kb) {
        // Round K over the bits: D folded by r into out[0]; the last such
        // round also writes h_K as elements into out[1].
        h->r.push_back(r);
        if (!h->rtab.p) TRY(h->rtab.alloc(kConstTable4Entries * sizeof(uint4)));
        hipLaunchKernelGGL(bitz::k_const_table4_1, dim3(1), dim3(1024), 0, s, r, static_cast<uint4*>(h->rtab.p));
        bitz::InnerBitsDesc A{};
        A.d = h->cur[0];
        A.bits = h->bits;
        A.dout = out[0];
        A.wout = K == h->kb ? out[1] : nullptr;
        A.pairs = ((h->live + 1) / 2 + 1) / 2;  // over the live prefix
        A.live = h->live;
        A.rtab4 = static_cast<const uint4*>(h->rtab.p);
        A.acc = h->bacc.u32() + bitz::kInnerBitsAccWords * K;
        A.blocks = bits_grid(A.pairs);
        if (A.wout) monomials(h->r.data(), K, A.mono);
        h->tm.begin(s, "inner_fold_next", 24.0 * (double)h->len + (A.wout ? 8.0 * (double)h->len : 0.0));
        TRY(launch_inner_bits(K, A, A.blocks, s));
        h->tm.end(s);
        HIP_TRY(hipMemcpyAsync(h->host.p, A.acc, bitz::kInnerBitsAccWords * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        h->tm.collect();
        bits_round_coeffs(static_cast<const uint32_t*>(h->host.p), K, h->r.data(), co);
        h->cur[0] = out[0];
        h->cur[1] = A.wout ? out[1] : nullptr;
        if (A.wout) h->bits = nullptr;  // h_kb is in cur[1]: the general kernels from here
        h->where = (h->where == 1) ? 2 : 1;
        h->len = next_len;
        h->live = (h->live + 1) / 2;
        h->round += 1;
        return BITZ_OK;
    }
    if (next_len == 1) {
        hipLaunchKernelGGL(bitz::k_fold, dim3(1), dim3(kBlock), 0, s, h->cur[0], out[0], (size_t)1, r);
        hipLaunchKernelGGL(bitz::k_fold, dim3(1), dim3(kBlock), 0, s, h->cur[1], out[1], (size_t)1, r);
    } else {
        // r's 4-bit table, then the fold with lane-table products.
        if (!h->rtab.p) TRY(h->rtab.alloc(kConstTable4Entries * sizeof(uint4)));
        hipLaunchKernelGGL(bitz::k_const_table4_1, dim3(1), dim3(1024), 0, s, r, static_cast<uint4*>(h->rtab.p));
        const size_t out_pairs = ((h->live + 1) / 2 + 1) / 2;  // over the live prefix
        h->tm.begin(s, "inner_fold_next", 48.0 * (double)h->len);
        hipLaunchKernelGGL(bitz::k_inner_fold_lane, dim3(inner_grid(out_pairs)), dim3(bitz::kInnerBlock),
                           bitz::kInnerGridSmem, s, h->cur[0], h->cur[1], out[0], out[1], out_pairs, h->live,
                           static_cast<const uint4*>(h->rtab.p), h->acc.u32() + 8 * (h->round + 1));
        h->tm.end(s);
        has_next = true;
    }
    TRY(check_launch());
    h->cur[0] = out[0];
    h->cur[1] = out[1];
    h->where = (h->where == 1) ? 2 : 1;
    h->len = next_len;
    h->live = (h->live + 1) / 2;
    h->round += 1;
    if (has_next) return inner_read(h, h->round, co);
    memset(co, 0, 32);
    return BITZ_OK;
}

extern "C" int32_t bitz_inner_finish(bitz_inner* h, bitz_f128* evals) {
    if (!h || !evals) return fail(BITZ_ERR_ARG, "null pointer");
    if (h->round != h->n) return fail(BITZ_ERR_STATE, "rounds remain");
    HIP_TRY(hipMemcpyAsync(evals, h->cur[0], sizeof(f128), hipMemcpyDeviceToHost, h->so.s));
    HIP_TRY(hipMemcpyAsync(evals + 1, h->cur[1], sizeof(f128), hipMemcpyDeviceToHost, h->so.s));
    HIP_TRY(hipStreamSynchronize(h->so.s));
    return BITZ_OK;
}

extern "C" void bitz_inner_free(bitz_inner* h) {
    if (h) {
        (void)hipStreamSynchronize(h->so.s);
        delete h;
    }
}

// The transcript-driven inner sumcheck through the per-round handle (begun
// by the caller, freed here).
static int32_t prove_inner_rounds(bitz_inner* h, bitz_transcript* t, bitz_f128 initial_claim, uint32_t n,
                                  bitz_f128* round_polys, bitz_f128* eval_points, bitz_f128* final_claim,
                                  bitz_f128* evals) {
    int32_t rc = BITZ_OK;
    bitz_f128 co[2];
    f128 claim = F(initial_claim);
    if ((rc = bitz_inner_round0(h, co))) goto done;
    for (uint32_t k = 0; k < n; ++k) {
        const f128 r = recover_round(t->t, claim, co, 2, round_polys + 3 * k, eval_points + k);
        if ((rc = bitz_inner_fold_next(h, B(r), co))) goto done;
    }
    if ((rc = bitz_inner_finish(h, evals))) goto done;
    *final_claim = B(claim);
done:
    bitz_inner_free(h);
    return rc;
}

static int32_t prove_inner_impl(bitz_transcript* t, bitz_f128 initial_claim, const bitz_f128* matrix,
                                uint32_t matrix_vars, const bitz_f128* witness, uint32_t witness_vars, uint32_t flags,
                                bitz_f128* round_polys, bitz_f128* eval_points, bitz_f128* final_claim,
                                bitz_f128* evals) {
    if (!t || !final_claim || !evals || (matrix_vars && (!round_polys || !eval_points)))
        return fail(BITZ_ERR_ARG, "null pointer");
    if (witness_vars != matrix_vars)
        return fail(BITZ_ERR_PRODUCT_DIMS, "matrix and witness have different num_vars");
    bitz_inner* h = nullptr;
    TRY(bitz_inner_begin(&h, matrix, matrix_vars, witness, witness_vars, flags));
    return prove_inner_rounds(h, t, initial_claim, matrix_vars, round_polys, eval_points, final_claim, evals);
}

extern "C" int32_t bitz_prove_inner_sumcheck(bitz_transcript* t, bitz_f128 initial_claim,
                                             const bitz_f128* matrix, uint32_t matrix_vars,
                                             const bitz_f128* witness, uint32_t witness_vars,
                                             uint32_t flags, bitz_f128* round_polys,
                                             bitz_f128* eval_points, bitz_f128* final_claim,
                                             bitz_f128* evals) {
    return prove_inner_impl(t, initial_claim, matrix, matrix_vars, witness, witness_vars, flags, round_polys,
                            eval_points, final_claim, evals);
}

extern "C" int32_t bitz_prove_inner_sumcheck_packed(bitz_transcript* t, bitz_f128 initial_claim,
                                                    const bitz_f128* matrix, uint32_t matrix_vars,
                                                    const uint64_t* assignment_bits, uint32_t assignment_vars,
                                                    uint32_t flags, bitz_f128* round_polys, bitz_f128* eval_points,
                                                    bitz_f128* final_claim, bitz_f128* evals) {
    if (!t || !final_claim || !evals || (matrix_vars && (!round_polys || !eval_points)))
        return fail(BITZ_ERR_ARG, "null pointer");
    bitz_inner* h = nullptr;
    TRY(bitz_inner_begin_packed(&h, matrix, matrix_vars, assignment_bits, assignment_vars, flags));
    return prove_inner_rounds(h, t, initial_claim, matrix_vars, round_polys, eval_points, final_claim, evals);
}

// Device buffers of the queued inner prover: the inputs, the ping-pong
// tables P (2 x capP entries: D then h) and Q (2 x capQ), n challenge tables,
// the n challenges and the round sums (8 words per round, then D(r), h(r);
// (n + 2) x 8 words, zero on entry).
struct InnerQueued {
    const f128* d = nullptr;
    const f128* w = nullptr;
    f128* P = nullptr;
    f128* Q = nullptr;
    size_t capP = 0, capQ = 0;
    uint4* rt4 = nullptr;
    const f128* chal = nullptr;
    uint32_t* acc = nullptr;
};

// Enqueues every round of the queued inner sumcheck over 2^n entries (n >= 1)
// on stream s: the challenge tables, round 0 and the grid-wide fold rounds
// (lane-table products, r's 4-bit table) back to back, the middle rounds in
// the chunk stage, the last rounds in one workgroup.  ts (k0 + 2 entries,
// or null) times the launches when profiling is on.
static int32_t inner_queued_enqueue(const InnerQueued& B, int n, hipStream_t s, std::vector<Timed>* ts) {
    hipLaunchKernelGGL(bitz::k_const_tables4, dim3((unsigned)n), dim3(512), 0, s, B.chal, B.rt4);
    TRY(check_launch());
    inner_set_attributes();
    const uint4* rt = B.rt4;
    const int k0 = std::max(0, n - bitz::kInnerTailMaxLog);
    // Rounds [kc, k0) run in the chunk stage (each workgroup folds its chunk
    // of 2^kInnerChunkLog inputs round after round in LDS), the grid-wide
    // kernels take [0, kc).
    const int kc = k0 - std::max(0, std::min(bitz::kInnerChunkRounds, k0 - 1));
    auto grid = [](size_t work) { return inner_grid(work); };
    auto timed = [&](int i) -> Timed* { return ts ? &(*ts)[i] : nullptr; };
    const size_t N = (size_t)1 << n;
    size_t len = N;
    const f128* cur[2] = {B.d, B.w};
    int where = 0;
    // The grid-wide fold rounds run back to back: one event pair spans them.
    double fold_bytes = 0;
    for (int k = 1; k < kc; ++k) fold_bytes += 96.0 * (double)(N >> k);
    for (int k = 0; k < kc; ++k) {
        if (k == 0) {
            if (Timed* t = timed(0)) t->begin(s, "inner_round0_lane", 32.0 * (double)len);
            hipLaunchKernelGGL(bitz::k_inner_r0_lane, dim3(grid(len / 2)), dim3(bitz::kInnerBlock), bitz::kInnerGridSmem,
                               s, cur[0], cur[1], len / 2, B.acc, nullptr);
        } else {
            f128* dst = (where == 1) ? B.Q : B.P;
            const size_t cap = (where == 1) ? B.capQ : B.capP;
            f128* out[2] = {dst, dst + cap};
            const size_t out_pairs = len / 4;
            if (k == 1)
                if (Timed* t = timed(1)) {
                    t->begin(s, "inner_fold_lane", fold_bytes);
                    t->launches = (uint64_t)(kc - 1);
                }
            hipLaunchKernelGGL(bitz::k_inner_fold_lane, dim3(grid(out_pairs)), dim3(bitz::kInnerBlock),
                               bitz::kInnerGridSmem, s, cur[0], cur[1], out[0], out[1], out_pairs, len,
                               rt + (size_t)(k - 1) * kConstTable4Entries, B.acc + 8 * (size_t)k);
            cur[0] = out[0];
            cur[1] = out[1];
            where = (where == 1) ? 2 : 1;
            len /= 2;
        }
        if (k == 0 || k == kc - 1)
            if (Timed* t = timed(k == 0 ? 0 : 1)) t->end(s);
        TRY(check_launch());
    }
    if (kc < k0) {
        // len = 2^(n - kc + 1) entries per table before the fold by r_{kc-1}
        f128* dst = (where == 1) ? B.Q : B.P;
        const size_t cap = (where == 1) ? B.capQ : B.capP;
        bitz::InnerTail ca{};
        ca.d = cur[0];
        ca.w = cur[1];
        ca.rtabs4 = rt;
        ca.acc = B.acc;
        ca.dout = dst;
        ca.wout = dst + cap;
        ca.n = n;
        ca.k0 = kc;
        ca.fold_first = 1;
        ca.lg_in = bitz::kInnerChunkLog;
        ca.s = k0 - kc;
        const unsigned chunks = (unsigned)(len >> bitz::kInnerChunkLog);
        if (Timed* t = timed(k0 + 1)) t->begin(s, "inner_chunk", 0.0);
        hipLaunchKernelGGL(bitz::k_inner_tail, dim3(chunks), dim3(bitz::kStageBlock), bitz::kInnerTailSmem, s, ca);
        if (Timed* t = timed(k0 + 1)) t->end(s);
        TRY(check_launch());
        cur[0] = ca.dout;
        cur[1] = ca.wout;
        where = (where == 1) ? 2 : 1;
        len >>= (k0 - kc + 1);
    }
    bitz::InnerTail ta{};
    ta.d = cur[0];
    ta.w = cur[1];
    ta.rtabs4 = rt;
    ta.acc = B.acc;
    ta.fin = reinterpret_cast<f128*>(B.acc + 8 * (size_t)n);
    ta.n = n;
    ta.k0 = k0;
    ta.fold_first = kc == k0 && k0 > 0;
    ta.lg_in = ta.fold_first ? n - k0 + 1 : n - k0;
    ta.s = n - k0;
    if (Timed* t = timed(k0)) t->begin(s, "inner_tail", 0.0);
    hipLaunchKernelGGL(bitz::k_inner_tail, dim3(1), dim3(bitz::kStageBlock), bitz::kInnerTailSmem, s, ta);
    if (Timed* t = timed(k0)) t->end(s);
    return check_launch();
}

// The round polynomials from the queued prover's sums w ((n + 1) x 8 words):
// c1 = claim + c2 (g(0) + g(1) = claim in characteristic 2,
// sumcheck.rs:760-781), claim <- g(r_k), as the transcript-driven prover.
static void inner_queued_recover(const uint32_t* w, int n, const bitz_f128* challenges, f128 claim,
                                 bitz_f128* round_polys, bitz_f128* eval_points, bitz_f128* final_claim,
                                 bitz_f128* evals) {
    for (int k = 0; k < n; ++k) {
        f128 c[3];
        memcpy(&c[0], w + 8 * k, 16);
        memcpy(&c[2], w + 8 * k + 4, 16);
        c[1] = hostf::add(claim, c[2]);
        const f128 r = F(challenges[k]);
        claim = hostf::horner(c, 3, r);
        for (int t = 0; t < 3; ++t) round_polys[3 * k + t] = B(c[t]);
        eval_points[k] = challenges[k];
    }
    memcpy(evals, w + 8 * n, 32);
    *final_claim = B(claim);
}

// The inner sumcheck with the challenges known in advance: every round
// queued back to back (inner_queued_enqueue), one read of every round's
// [c0, c2] and of D(r), h(r) at the end.
extern "C" int32_t bitz_prove_inner_sumcheck_with_challenges(const bitz_f128* challenges, bitz_f128 initial_claim,
                                                             const bitz_f128* matrix, uint32_t matrix_vars,
                                                             const bitz_f128* witness, uint32_t witness_vars,
                                                             uint32_t flags, bitz_f128* round_polys,
                                                             bitz_f128* eval_points, bitz_f128* final_claim,
                                                             bitz_f128* evals) {
    if (!final_claim || !evals || (matrix_vars && (!round_polys || !eval_points || !challenges)))
        return fail(BITZ_ERR_ARG, "null pointer");
    if (witness_vars != matrix_vars)
        return fail(BITZ_ERR_PRODUCT_DIMS, "matrix and witness have different num_vars");
    bitz_inner* h = nullptr;
    TRY(bitz_inner_begin(&h, matrix, matrix_vars, witness, witness_vars, flags));
    std::unique_ptr<bitz_inner, void (*)(bitz_inner*)> guard(h, bitz_inner_free);
    const int n = h->n;
    hipStream_t s = h->so.s;
    f128 claim = F(initial_claim);
    if (n == 0) {
        HIP_TRY(hipMemcpyAsync(h->host.p, h->cur[0], sizeof(f128), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(static_cast<char*>(h->host.p) + 16, h->cur[1], sizeof(f128), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        memcpy(evals, h->host.p, 32);
        *final_claim = B(claim);
        return BITZ_OK;
    }
    Dev chal, rt4;
    TRY(chal.alloc((size_t)n * sizeof(f128)));
    TRY(rt4.alloc((size_t)n * kConstTable4Entries * sizeof(uint4)));
    Pinned up;
    TRY(up.alloc((size_t)n * sizeof(f128)));
    memcpy(up.p, challenges, (size_t)n * sizeof(f128));
    HIP_TRY(hipMemcpyAsync(chal.p, up.p, (size_t)n * sizeof(f128), hipMemcpyHostToDevice, s));
    InnerQueued q;
    q.d = h->cur[0];
    q.w = h->cur[1];
    q.P = h->P.f();
    q.Q = h->Q.f();
    q.capP = h->capP;
    q.capQ = h->capQ;
    q.rt4 = static_cast<uint4*>(rt4.p);
    q.chal = chal.f();
    q.acc = h->acc.u32();
    std::vector<Timed> ts(std::max(0, n - bitz::kInnerTailMaxLog) + 2);
    TRY(inner_queued_enqueue(q, n, s, &ts));
    Pinned down;
    TRY(down.alloc(((size_t)n + 1) * 32));
    HIP_TRY(hipMemcpyAsync(down.p, h->acc.p, ((size_t)n + 1) * 32, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (auto& t : ts) t.collect();
    inner_queued_recover(static_cast<const uint32_t*>(down.p), n, challenges, claim, round_polys, eval_points,
                         final_claim, evals);
    return BITZ_OK;
}

// The queued inner prover without any allocation: the ping-pong tables,
// challenge tables, challenges and round sums live in the device's static
// scratch (2^20 entries per table at most).  Calls that overlap on one GPU
// share that scratch and race; the round sums are cleared by a kernel at the
// start of every call.
extern "C" int32_t bitz_prove_inner_sumcheck_with_challenges_noalloc(
    const bitz_f128* challenges, bitz_f128 initial_claim, const bitz_f128* matrix, uint32_t matrix_vars,
    const bitz_f128* witness, uint32_t witness_vars, void* stream, bitz_f128* round_polys, bitz_f128* eval_points,
    bitz_f128* final_claim, bitz_f128* evals) {
    if (!matrix || !witness || !final_claim || !evals ||
        (matrix_vars && (!round_polys || !eval_points || !challenges)))
        return fail(BITZ_ERR_ARG, "null pointer");
    if (witness_vars != matrix_vars)
        return fail(BITZ_ERR_PRODUCT_DIMS, "matrix and witness have different num_vars");
    if (matrix_vars > (uint32_t)kNoAllocMaxVars)
        return fail(BITZ_ERR_ARG, "num_vars " + std::to_string(matrix_vars) + " exceeds the static scratch's maximum of " +
                                      std::to_string(kNoAllocMaxVars));
    void* base = nullptr;
    TRY(scratch_address(&base));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int n = (int)matrix_vars;
    const f128* d = reinterpret_cast<const f128*>(matrix);
    const f128* w = reinterpret_cast<const f128*>(witness);
    f128 claim = F(initial_claim);
    if (n == 0) {
        HIP_TRY(hipMemcpyAsync(evals, d, sizeof(f128), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(evals + 1, w, sizeof(f128), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        *final_claim = B(claim);
        return BITZ_OK;
    }
    char* sc = static_cast<char*>(base);
    const size_t N = (size_t)1 << n;
    InnerQueued q;
    q.d = d;
    q.w = w;
    q.capP = std::max<size_t>(N / 2, 1);
    q.capQ = std::max<size_t>(N / 4, 1);
    q.P = reinterpret_cast<f128*>(sc);
    q.Q = reinterpret_cast<f128*>(sc + kNoAllocQOff);
    q.rt4 = reinterpret_cast<uint4*>(sc + kNoAllocTabOff);
    f128* chal = reinterpret_cast<f128*>(sc + kNoAllocChalOff);
    q.chal = chal;
    q.acc = reinterpret_cast<uint32_t*>(sc + kNoAllocAccOff);
    const size_t acc_words = ((size_t)n + 2) * 8;
    hipLaunchKernelGGL(bitz::k_zero_words, dim3(1), dim3(256), 0, s, q.acc, acc_words);
    TRY(check_launch());
    HIP_TRY(hipMemcpyAsync(chal, challenges, (size_t)n * sizeof(f128), hipMemcpyHostToDevice, s));
    TRY(inner_queued_enqueue(q, n, s, nullptr));
    alignas(16) uint32_t sums[(kNoAllocMaxVars + 1) * 8];
    HIP_TRY(hipMemcpyAsync(sums, q.acc, ((size_t)n + 1) * 32, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    inner_queued_recover(sums, n, challenges, claim, round_polys, eval_points, final_claim, evals);
    return BITZ_OK;
}

// ============================================================ post-GKR =====
// prove_factors (crates/post_gkr/src/lib.rs:125-148) with the degree-two
// rounds of crates/post_gkr/src/sumcheck.rs:99-178: the claim
// <rows (x) columns, f> = target over the committed bits f (2^t x 2^s, bit
// (c << t) | b, packed 128 to an element, t >= 7) becomes MLE[f](point) =
// evaluation.  Round 0 runs off the packed bits (k_pg_first: XOR sums of
// row weights, one product per element); the device then writes W = the rows
// folded at r_0 tensored with the columns and V = the bits bound at r_0
// (k_pg_expand), and the other t + s - 1 rounds are the inner sumcheck's
// rounds over (W, V): the same [sum w0 v0, sum (w0 + w1)(v0 + v1)] and fold.
// Each round: its [a0, a2] as a prover message, the grind, the challenge;
// the closing evaluation as a prover message.
extern "C" int32_t bitz_post_gkr_prove(bitz_transcript* t, const bitz_f128* rows, uint32_t log_rows,
                                       const bitz_f128* columns, uint32_t log_columns, bitz_f128 target,
                                       const bitz_f128* packed, size_t packed_len, uint32_t security_bits,
                                       uint32_t flags, bitz_f128* point, bitz_f128* evaluation) {
    if (!t || !rows || !columns || !packed || !point || !evaluation) return fail(BITZ_ERR_ARG, "null pointer");
    if (log_rows < 7 || log_rows + log_columns > 40)
        return fail(BITZ_ERR_ARG, "the claim is over 2^t x 2^s bits with 7 <= t, t + s <= 40");
    if (security_bits != 100 && security_bits != 128)
        return fail(BITZ_ERR_ARG, "security_bits is a SecurityLevel: 100 or 128");
    const size_t per_column = (size_t)1 << (log_rows - 7), ncols = (size_t)1 << log_columns;
    if (packed_len != per_column * ncols)  // lib.rs:133-135
        return fail(BITZ_ERR_WITNESS_LENGTH, "the packed witness is not one element per 128 bits of the claim");
    StreamOwner so;
    TRY(so.create());
    hipStream_t s = so.s;
    const size_t nrows = (size_t)1 << log_rows;
    Dev drows, dcols, dpacked;
    const f128* R = reinterpret_cast<const f128*>(rows);
    const f128* Cw = reinterpret_cast<const f128*>(columns);
    const f128* Pk = reinterpret_cast<const f128*>(packed);
    if (!(flags & BITZ_DEVICE)) {
        TRY(upload(drows, rows, nrows, flags, s));
        TRY(upload(dcols, columns, ncols, flags, s));
        TRY(upload(dpacked, packed, packed_len, flags, s));
        R = drows.f();
        Cw = dcols.f();
        Pk = dpacked.f();
    }
    Dev acc;
    TRY(acc.alloc(64));
    HIP_TRY(hipMemsetAsync(acc.p, 0, 64, s));
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((packed_len + 3) / 4, 2048));
    hipLaunchKernelGGL(bitz::k_pg_first, dim3(blocks), dim3(256), 0, s, R, (int)log_rows, Cw, Pk, packed_len,
                       acc.u32());
    TRY(check_launch());
    Pinned host;
    TRY(host.alloc(64));
    HIP_TRY(hipMemcpyAsync(host.p, acc.p, 48, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    f128 sums[3];
    memcpy(sums, host.p, 48);
    if (!hostf::eq(sums[0], F(target)))  // lib.rs:136-138
        return fail(BITZ_ERR_CLAIM, "the claim's weights against the witness do not give the target");
    static const uint8_t kLabel[] = "bitz/pcs/sumcheck/v1";  // lib.rs:71
    const uint32_t gbits = bitz::grinding_bits(security_bits, 2);
    // lib.rs:286-289 advance: h' = a0 + r (h + a2 + r a2).
    auto advance = [](f128 h, f128 a0, f128 a2, f128 r) {
        return hostf::add(a0, hostf::mul(r, hostf::add(hostf::add(h, a2), hostf::mul(r, a2))));
    };
    f128 msg[2] = {sums[1], sums[2]};
    t->t.absorb_elems(1, msg, 2);
    t->t.grind(kLabel, sizeof kLabel - 1, gbits);
    f128 r = t->t.squeeze();
    point[0] = B(r);
    f128 h = advance(F(target), msg[0], msg[1], r);
    const int m1 = (int)(log_rows + log_columns) - 1;
    const size_t n = (size_t)1 << m1, half_rows = nrows / 2;
    Dev rf, W, V;
    TRY(rf.alloc(half_rows * sizeof(f128)));
    TRY(W.alloc(n * sizeof(f128)));
    TRY(V.alloc(n * sizeof(f128)));
    hipLaunchKernelGGL(bitz::k_fold, dim3(grid_for(half_rows)), dim3(kBlock), 0, s, R, rf.f(), half_rows, r);
    hipLaunchKernelGGL(bitz::k_pg_expand, dim3(grid_for(n)), dim3(kBlock), 0, s, rf.f(), (int)log_rows, Cw, Pk, r,
                       W.f(), V.f(), n);
    TRY(check_launch());
    HIP_TRY(hipStreamSynchronize(s));  // the inner handle works on a stream of its own
    bitz_inner* ih = nullptr;
    TRY(bitz_inner_begin(&ih, reinterpret_cast<const bitz_f128*>(W.p), (uint32_t)m1,
                         reinterpret_cast<const bitz_f128*>(V.p), (uint32_t)m1, BITZ_DEVICE));
    std::unique_ptr<bitz_inner, void (*)(bitz_inner*)> guard(ih, bitz_inner_free);
    bitz_f128 co[2];
    TRY(bitz_inner_round0(ih, co));
    for (int k = 1; k <= m1; ++k) {
        const f128 m[2] = {F(co[0]), F(co[1])};
        t->t.absorb_elems(1, m, 2);
        t->t.grind(kLabel, sizeof kLabel - 1, gbits);
        r = t->t.squeeze();
        point[k] = B(r);
        h = advance(h, m[0], m[1], r);
        TRY(bitz_inner_fold_next(ih, B(r), co));
    }
    bitz_f128 ev[2];
    TRY(bitz_inner_finish(ih, ev));
    const f128 v = F(ev[1]);
    t->t.absorb_elems(1, &v, 1);  // sumcheck.rs:172-177
    *evaluation = ev[1];
    return BITZ_OK;
}

// ========================================================= matrices =======
struct bitz_matrices {
    size_t rows = 0, rows_padded = 0, cols = 0;
    int row_vars = 0, col_vars = 0;
    Dev col_starts, row_idx, which, values, eq_rows, row_scale, scales, point, scratch;
    Dev rs[3], cs[3], vs[3];  // row-major copies of A, B, C (products)
    Dev prod;                 // bind_and_batch: one product per nonzero
    Dev long_chunks;          // {column, first, end} runs of the long columns
    size_t nnz = 0, n_long_chunks = 0;
    std::mutex mu;            // calls on one handle share its scratch and stream: one at a time
    // The last bind's work (on whichever stream ran it): the next bind, on
    // any stream, waits for it before it writes the shared scratch.
    hipEvent_t bind_done = nullptr;
    StreamOwner so;  // last: destroyed (and synchronised) before the buffers
    ~bitz_matrices() {
        if (bind_done) {
            (void)hipEventSynchronize(bind_done);
            (void)hipEventDestroy(bind_done);
        }
    }
};

// Device bytes a prepared matrices handle holds (every buffer as allocated,
// size classes included).
static uint64_t matrices_device_bytes(const bitz_matrices* h) {
    uint64_t b = h->col_starts.bytes + h->row_idx.bytes + h->which.bytes + h->values.bytes + h->eq_rows.bytes +
                 h->row_scale.bytes + h->scales.bytes + h->point.bytes + h->scratch.bytes + h->prod.bytes +
                 h->long_chunks.bytes;
    for (int k = 0; k < 3; ++k) b += h->rs[k].bytes + h->cs[k].bytes + h->vs[k].bytes;
    return b;
}

extern "C" int32_t bitz_matrices_device_bytes(const bitz_matrices* h, uint64_t* bytes) {
    if (!h || !bytes) return fail(BITZ_ERR_ARG, "null pointer");
    *bytes = matrices_device_bytes(h);
    return BITZ_OK;
}

extern "C" int32_t bitz_matrices_prepare(bitz_matrices** out, size_t rows, uint32_t row_vars,
                                         uint32_t col_vars, const uint64_t* const row_starts[3],
                                         const uint32_t* const columns[3],
                                         const bitz_f128* const values[3]) {
    if (!out || !row_starts || !columns || !values) return fail(BITZ_ERR_ARG, "null pointer");
    if (row_vars > 30 || col_vars > 31 || rows > ((size_t)1 << row_vars))
        return fail(BITZ_ERR_ARG, "matrix shape out of range");
    const size_t cols = (size_t)1 << col_vars;
    // The device keeps only entries whose value is nonzero: a zero entry adds
    // nothing to A h, B h, C h or to D (a projection can map coefficients to
    // zero, and the layout keeps their positions).
    auto live = [&](int m, uint64_t k) { return values[m][k].lo != 0 || values[m][k].hi != 0; };
    // Column-major copy of all three matrices: counting sort by column.
    std::vector<uint32_t> count(cols + 1, 0);
    size_t nnz = 0;
    for (int m = 0; m < 3; ++m) {
        if (!row_starts[m]) return fail(BITZ_ERR_ARG, "null row_starts");
        const size_t mn = row_starts[m][rows];
        for (size_t k = 0; k < mn; ++k) {
            if (columns[m][k] >= cols) return fail(BITZ_ERR_ARG, "column index out of range");
            if (live(m, k)) {
                count[columns[m][k] + 1]++;
                ++nnz;
            }
        }
    }
    if (nnz >= 0xFFFFFFFFull) return fail(BITZ_ERR_ARG, "too many nonzeros");
    for (size_t j = 0; j < cols; ++j) count[j + 1] += count[j];
    std::vector<uint32_t> starts(count), rowv(nnz);
    std::vector<uint8_t> whichv(nnz);
    std::vector<bitz_f128> vals(std::max<size_t>(nnz, 1));
    std::vector<uint32_t> fill(count.begin(), count.end() - 1);
    for (int m = 0; m < 3; ++m)
        for (size_t i = 0; i < rows; ++i)
            for (uint64_t k = row_starts[m][i]; k < row_starts[m][i + 1]; ++k) {
                if (!live(m, k)) continue;
                const uint32_t pos = fill[columns[m][k]]++;
                rowv[pos] = (uint32_t)i;
                whichv[pos] = (uint8_t)m;
                vals[pos] = values[m][k];
            }
    // Columns longer than kBindLongColumn entries are summed by workgroups
    // over runs of at most kBindLongRun entries instead of by one thread.
    std::vector<uint32_t> chunks;
    for (size_t j = 0; j < cols; ++j)
        if (starts[j + 1] - starts[j] > bitz::kBindLongColumn)
            for (uint32_t lo = starts[j]; lo < starts[j + 1]; lo += bitz::kBindLongRun) {
                chunks.push_back((uint32_t)j);
                chunks.push_back(lo);
                chunks.push_back(std::min<uint32_t>(lo + bitz::kBindLongRun, starts[j + 1]));
            }
    auto* h = new bitz_matrices;
    auto cleanup = [&](int32_t c) {
        delete h;
        return c;
    };
    int32_t rc;
    if ((rc = h->so.create())) return cleanup(rc);
    h->rows = rows;
    h->rows_padded = (size_t)1 << row_vars;
    h->cols = cols;
    h->row_vars = (int)row_vars;
    h->col_vars = (int)col_vars;
    if ((rc = h->col_starts.alloc(starts.size() * 4))) return cleanup(rc);
    if ((rc = h->row_idx.alloc(std::max<size_t>(nnz, 1) * 4))) return cleanup(rc);
    if ((rc = h->which.alloc(std::max<size_t>(nnz, 1)))) return cleanup(rc);
    if ((rc = h->values.alloc(vals.size() * sizeof(f128)))) return cleanup(rc);
    h->nnz = nnz;
    if ((rc = h->prod.alloc(std::max<size_t>(nnz, 1) * sizeof(f128)))) return cleanup(rc);
    if ((rc = h->eq_rows.alloc(h->rows_padded * sizeof(f128)))) return cleanup(rc);
    if ((rc = h->row_scale.alloc(3 * h->rows_padded * sizeof(f128)))) return cleanup(rc);
    if ((rc = h->scales.alloc(3 * sizeof(f128)))) return cleanup(rc);
    if ((rc = h->point.alloc(std::max<uint32_t>(row_vars, 1) * sizeof(f128)))) return cleanup(rc);
    h->n_long_chunks = chunks.size() / 3;
    if ((rc = h->long_chunks.alloc(std::max<size_t>(chunks.size(), 1) * 4))) return cleanup(rc);
    if ((!chunks.empty() &&
         hipMemcpy(h->long_chunks.p, chunks.data(), chunks.size() * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(h->col_starts.p, starts.data(), starts.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        (nnz && hipMemcpy(h->row_idx.p, rowv.data(), nnz * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (nnz && hipMemcpy(h->which.p, whichv.data(), nnz, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(h->values.p, vals.data(), vals.size() * sizeof(f128), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(BITZ_ERR_DEVICE, "matrix upload failed"));
    for (int m = 0; m < 3; ++m) {
        std::vector<uint32_t> rs32(rows + 1), cs32;
        std::vector<bitz_f128> vs;
        rs32[0] = 0;
        for (size_t i = 0; i < rows; ++i) {
            for (uint64_t k = row_starts[m][i]; k < row_starts[m][i + 1]; ++k)
                if (live(m, k)) {
                    cs32.push_back(columns[m][k]);
                    vs.push_back(values[m][k]);
                }
            rs32[i + 1] = (uint32_t)cs32.size();
        }
        const size_t mn = cs32.size();
        if ((rc = h->rs[m].alloc((rows + 1) * 4)) || (rc = h->cs[m].alloc(std::max<size_t>(mn, 1) * 4)) ||
            (rc = h->vs[m].alloc(std::max<size_t>(mn, 1) * sizeof(f128))))
            return cleanup(rc);
        if (hipMemcpy(h->rs[m].p, rs32.data(), (rows + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
            (mn && hipMemcpy(h->cs[m].p, cs32.data(), mn * 4, hipMemcpyHostToDevice) != hipSuccess) ||
            (mn && hipMemcpy(h->vs[m].p, vs.data(), mn * sizeof(f128), hipMemcpyHostToDevice) != hipSuccess))
            return cleanup(fail(BITZ_ERR_DEVICE, "matrix upload failed"));
    }
    *out = h;
    return BITZ_OK;
}

// Enqueues D = bind_and_batch(r_x, rho) into target (2^col_vars, device) on
// stream s, after copying the row point (row_vars entries, host or device)
// into the handle.  Uses the handle's scratch: callers hold h->mu while
// enqueueing, and the bind_done event orders the kernels across streams.
// A bind's working buffers (row point, eq table, batching scales, one product
// per nonzero) for a caller that keeps its own set: the matrices handle's
// set serves every other bind, one at a time.
struct BindScratch {
    Dev point, eq_rows, eq_scratch, row_scale, prod;
    size_t rows_padded = 0, nnz = 0;
    int row_vars = -1;
    int32_t ensure(const bitz_matrices* h) {
        if (prod.p && rows_padded == h->rows_padded && nnz == h->nnz && row_vars == h->row_vars) return BITZ_OK;
        rows_padded = h->rows_padded;
        nnz = h->nnz;
        row_vars = h->row_vars;
        TRY(point.alloc(std::max<int>(h->row_vars, 1) * sizeof(f128)));
        TRY(eq_rows.alloc(h->rows_padded * sizeof(f128)));
        TRY(row_scale.alloc(3 * h->rows_padded * sizeof(f128)));
        return prod.alloc(std::max<size_t>(h->nnz, 1) * sizeof(f128));
    }
};

static int32_t bind_enqueue(bitz_matrices* h, const f128* row_point, bool point_on_device, f128 rho, f128* target,
                            hipStream_t s, const f128* rho_dev = nullptr, BindScratch* own = nullptr) {
    // The point, eq table, scales and products live in the handle unless the
    // caller brings its own: then wait for the previous bind (possibly
    // queued on another stream) first.
    if (!own && h->bind_done) HIP_TRY(hipStreamWaitEvent(s, h->bind_done, 0));
    if (own) TRY(own->ensure(h));
    f128* point = own ? own->point.f() : h->point.f();
    f128* eq_rows = own ? own->eq_rows.f() : h->eq_rows.f();
    f128* row_scale = own ? own->row_scale.f() : h->row_scale.f();
    f128* prod = own ? own->prod.f() : h->prod.f();
    // A caller with its own scratch and the point on the device (the PIOP
    // engine: the outer sumcheck's challenges, which outlive the bind) is
    // read in place; otherwise the point is copied first.
    const f128* pt = point;
    if (own && point_on_device)
        pt = row_point;
    else if (h->row_vars)
        HIP_TRY(hipMemcpyAsync(point, row_point, (size_t)h->row_vars * sizeof(f128),
                               point_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    // batching scales 1, rho, rho^2 (matrix.rs:533-545), by value or (rho_dev:
    // rho, rho^2) from the device; small row domains write the scales with
    // the eq table in one launch
    if (rho_dev && h->row_vars <= bitz::kEqRowsMaxVars) {
        eq_rows_attr();
        hipLaunchKernelGGL(bitz::k_eq_rows, dim3(bitz::eq_rows_grid(h->row_vars)), dim3(1024), bitz::kLaneSmem1024, s,
                           pt, h->row_vars, row_scale, rho_dev);
    } else {
        TRY(eq_build(pt, h->row_vars, eq_rows, own ? own->eq_scratch : h->scratch, s));
    }
    if (rho_dev && h->row_vars > bitz::kEqRowsMaxVars)
        hipLaunchKernelGGL(bitz::k_row_scales_dev, dim3(grid_for(h->rows_padded)), dim3(kBlock), 0, s, eq_rows,
                           rho_dev, row_scale, h->rows_padded);
    else if (!rho_dev)
        hipLaunchKernelGGL(bitz::k_row_scales, dim3(grid_for(h->rows_padded)), dim3(kBlock), 0, s, eq_rows, rho,
                           hostf::mul(rho, rho), row_scale, h->rows_padded);
    // every nonzero's product (a converged pass), then each column's XOR over its run
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_bind_products),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kBindSmem);
        attr = true;
    }
    if (h->nnz) {
        const unsigned pgrid = (unsigned)std::max<size_t>(
            1, std::min<size_t>((h->nnz + bitz::kStageBlock - 1) / bitz::kStageBlock, 4096));
        hipLaunchKernelGGL(bitz::k_bind_products, dim3(pgrid), dim3(bitz::kStageBlock), bitz::kBindSmem, s,
                           h->row_idx.u32(), static_cast<const uint8_t*>(h->which.p), h->values.f(), row_scale,
                           h->rows_padded, prod, h->nnz);
    }
    const unsigned cgrid = (unsigned)std::max<size_t>(1, std::min<size_t>((h->cols + kBlock - 1) / kBlock, 1u << 20));
    hipLaunchKernelGGL(bitz::k_bind_segments, dim3(cgrid), dim3(kBlock), 0, s, h->col_starts.u32(), prod, target,
                       h->cols);
    if (h->n_long_chunks)  // the long columns, left zero above: XOR of each run's sum
        hipLaunchKernelGGL(bitz::k_bind_long, dim3((unsigned)h->n_long_chunks), dim3(kBlock), 0, s,
                           h->long_chunks.u32(), prod, target);
    TRY(check_launch());
    if (own) return BITZ_OK;
    if (!h->bind_done) HIP_TRY(hipEventCreateWithFlags(&h->bind_done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(h->bind_done, s));
    return BITZ_OK;
}

extern "C" int32_t bitz_bind_and_batch(bitz_matrices* h, const bitz_f128* row_point, uint32_t point_len,
                                       bitz_f128 rho, bitz_f128* out, uint32_t flags) {
    if (!h || !out || (point_len && !row_point)) return fail(BITZ_ERR_ARG, "null pointer");
    // matrix.rs:435-441: the row point must have num_row_vars coordinates.
    if (point_len != (uint32_t)h->row_vars) return fail(BITZ_ERR_ROW_POINT, "row point length");
    std::lock_guard<std::mutex> lock(h->mu);
    hipStream_t s = h->so.s;
    Dev dout;
    f128* target = reinterpret_cast<f128*>(out);
    if (!(flags & BITZ_DEVICE)) {
        TRY(dout.alloc(h->cols * sizeof(f128)));
        target = dout.f();
    }
    TRY(bind_enqueue(h, reinterpret_cast<const f128*>(row_point), (flags & BITZ_DEVICE) != 0, F(rho), target, s));
    if (!(flags & BITZ_DEVICE))
        HIP_TRY(hipMemcpyAsync(out, target, h->cols * sizeof(f128), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BITZ_OK;
}

extern "C" int32_t bitz_r1cs_products(bitz_matrices* h, const bitz_f128* assignment, bitz_f128* az, bitz_f128* bz,
                                      bitz_f128* cz, uint32_t flags) {
    if (!h || !assignment || !az || !bz || !cz) return fail(BITZ_ERR_ARG, "null pointer");
    std::lock_guard<std::mutex> lock(h->mu);
    hipStream_t s = h->so.s;
    Dev din, dout;
    const f128* hv = reinterpret_cast<const f128*>(assignment);
    if (!(flags & BITZ_DEVICE)) {
        TRY(upload(din, assignment, h->cols, 0, s));
        hv = din.f();
        TRY(dout.alloc(3 * h->rows_padded * sizeof(f128)));
    }
    bitz_f128* outs[3] = {az, bz, cz};
    for (int m = 0; m < 3; ++m) {
        f128* target = (flags & BITZ_DEVICE) ? reinterpret_cast<f128*>(outs[m]) : dout.f() + m * h->rows_padded;
        hipLaunchKernelGGL(bitz::k_r1cs_rows, dim3(grid_for(h->rows_padded)), dim3(kBlock), 0, s, h->rs[m].u32(),
                           h->cs[m].u32(), h->vs[m].f(), hv, h->rows, target, h->rows_padded);
        TRY(check_launch());
    }
    if (!(flags & BITZ_DEVICE))
        for (int m = 0; m < 3; ++m)
            HIP_TRY(hipMemcpyAsync(outs[m], dout.f() + m * h->rows_padded, h->rows_padded * sizeof(f128),
                                   hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BITZ_OK;
}

extern "C" int32_t bitz_r1cs_products_packed(bitz_matrices* h, const uint64_t* assignment_bits, bitz_f128* az,
                                             bitz_f128* bz, bitz_f128* cz, uint32_t flags) {
    if (!h || !assignment_bits || !az || !bz || !cz) return fail(BITZ_ERR_ARG, "null pointer");
    std::lock_guard<std::mutex> lock(h->mu);
    hipStream_t s = h->so.s;
    Dev din, dout;
    const uint64_t* bits = assignment_bits;
    if (!(flags & BITZ_DEVICE)) {
        const size_t words = (h->cols + 63) / 64;
        TRY(din.alloc(words * 8));
        HIP_TRY(hipMemcpyAsync(din.p, assignment_bits, words * 8, hipMemcpyHostToDevice, s));
        bits = static_cast<const uint64_t*>(din.p);
        TRY(dout.alloc(3 * h->rows_padded * sizeof(f128)));
    }
    bitz_f128* outs[3] = {az, bz, cz};
    for (int m = 0; m < 3; ++m) {
        f128* target = (flags & BITZ_DEVICE) ? reinterpret_cast<f128*>(outs[m]) : dout.f() + m * h->rows_padded;
        hipLaunchKernelGGL(bitz::k_r1cs_rows_bits, dim3(grid_for(h->rows_padded)), dim3(kBlock), 0, s,
                           h->rs[m].u32(), h->cs[m].u32(), h->vs[m].f(), bits, h->rows, target, h->rows_padded);
        TRY(check_launch());
    }
    if (!(flags & BITZ_DEVICE))
        for (int m = 0; m < 3; ++m)
            HIP_TRY(hipMemcpyAsync(outs[m], dout.f() + m * h->rows_padded, h->rows_padded * sizeof(f128),
                                   hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return BITZ_OK;
}

extern "C" int32_t bitz_assignment_mle(const uint64_t* words, size_t bit_len, size_t columns, bitz_f128* out,
                                       uint32_t flags) {
    if (!out || (bit_len && !words)) return fail(BITZ_ERR_ARG, "null pointer");
    if (bit_len != columns) return fail(BITZ_ERR_ASSIGNMENT_LEN, "assignment length differs from the columns");
    if (columns == 0 || !(words[0] & 1)) return fail(BITZ_ERR_ASSIGNMENT_CONST, "the first assignment bit must be one");
    size_t padded = 1;
    while (padded < columns) padded <<= 1;
    StreamOwner so;
    TRY(so.create());
    Dev dw, dout;
    const size_t nw = (bit_len + 63) / 64;
    TRY(dw.alloc(nw * 8));
    HIP_TRY(hipMemcpyAsync(dw.p, words, nw * 8, hipMemcpyHostToDevice, so.s));
    f128* target = reinterpret_cast<f128*>(out);
    if (!(flags & BITZ_DEVICE)) {
        TRY(dout.alloc(padded * sizeof(f128)));
        target = dout.f();
    }
    hipLaunchKernelGGL(bitz::k_assignment_bits, dim3(grid_for(padded)), dim3(kBlock), 0, so.s,
                       static_cast<const uint64_t*>(dw.p), bit_len, target, padded);
    TRY(check_launch());
    if (!(flags & BITZ_DEVICE))
        HIP_TRY(hipMemcpyAsync(out, target, padded * sizeof(f128), hipMemcpyDeviceToHost, so.s));
    HIP_TRY(hipStreamSynchronize(so.s));
    return BITZ_OK;
}

extern "C" int32_t bitz_assignment_pack(const uint64_t* words, size_t bit_len, size_t columns, uint64_t* out,
                                        uint32_t flags) {
    if (!out || (bit_len && !words)) return fail(BITZ_ERR_ARG, "null pointer");
    if (bit_len != columns) return fail(BITZ_ERR_ASSIGNMENT_LEN, "assignment length differs from the columns");
    if (columns == 0 || !(words[0] & 1)) return fail(BITZ_ERR_ASSIGNMENT_CONST, "the first assignment bit must be one");
    size_t padded = 1;
    while (padded < columns) padded <<= 1;
    const size_t nw = (padded + 63) / 64, used = (bit_len + 63) / 64;
    std::vector<uint64_t> packed(nw, 0);
    memcpy(packed.data(), words, used * 8);
    if (bit_len % 64) packed[used - 1] &= (~0ull) >> (64 - bit_len % 64);  // zero past the assignment
    if (!(flags & BITZ_DEVICE)) {
        memcpy(out, packed.data(), nw * 8);
        return BITZ_OK;
    }
    StreamOwner so;  // non-blocking: the copy does not wait for other streams
    TRY(so.create());
    HIP_TRY(hipMemcpyAsync(out, packed.data(), nw * 8, hipMemcpyHostToDevice, so.s));
    HIP_TRY(hipStreamSynchronize(so.s));
    return BITZ_OK;
}

extern "C" int32_t bitz_constraint_matrix_digest(size_t m_rows, size_t m_columns, const uint64_t* m_row_starts,
                                                 const uint32_t* m_positions, size_t rows, size_t columns,
                                                 const uint64_t* const row_starts[3], const uint32_t* const cols[3],
                                                 const bitz_f128* const values[3], uint8_t digest[32]) {
    if (!digest || !m_row_starts || !row_starts || !cols || !values) return fail(BITZ_ERR_ARG, "null pointer");
    for (int m = 0; m < 3; ++m)
        if (!row_starts[m]) return fail(BITZ_ERR_ARG, "null row_starts");
    // The byte stream is staged through a local buffer and hashed in large
    // pieces; u64 values are little-endian (x86 stores them so).
    bitz::Sha256 hs;
    std::vector<uint8_t> buf(1 << 16);
    size_t fill = 0;
    auto flush = [&] {
        hs.update(buf.data(), fill);
        fill = 0;
    };
    auto put = [&](const void* p, size_t n) {
        if (fill + n > buf.size()) flush();
        memcpy(buf.data() + fill, p, n);
        fill += n;
    };
    auto put_u64 = [&](uint64_t v) { put(&v, 8); };
    static const char kDomain[] = "bitz/spartan/constraint-matrices/v1";
    put(kDomain, sizeof kDomain - 1);
    put("M", 1);
    put_u64(m_rows);
    put_u64(m_columns);
    for (size_t i = 0; i < m_rows; ++i) {
        put_u64(m_row_starts[i + 1] - m_row_starts[i]);
        for (uint64_t k = m_row_starts[i]; k < m_row_starts[i + 1]; ++k) put_u64(m_positions[k]);
    }
    static const char* kLabel[3] = {"A", "B", "C"};
    for (int m = 0; m < 3; ++m) {
        put(kLabel[m], 1);
        put_u64(rows);
        put_u64(columns);
        for (size_t i = 0; i < rows; ++i) {
            put_u64(row_starts[m][i + 1] - row_starts[m][i]);
            for (uint64_t k = row_starts[m][i]; k < row_starts[m][i + 1]; ++k) {
                uint64_t entry[4] = {cols[m][k], 16, values[m][k].lo, values[m][k].hi};
                // column, then the coefficient's encoding: length 16 and
                // F128::to_bytes (crates/field/src/codec.rs:24-28)
                put(entry, sizeof entry);
            }
        }
    }
    flush();
    hs.finish(digest);
    return BITZ_OK;
}

// prove_spartan_piop over an assignment given as elements or (bits non-null)
// as its packed bits.
static int32_t piop_impl(bitz_transcript* t, bitz_matrices* h, const uint8_t digest[32], const bitz_f128* az,
                         const bitz_f128* bz, const bitz_f128* cz, uint32_t row_vars, const bitz_f128* assignment,
                         const uint64_t* bits, uint32_t col_vars, uint32_t flags, bitz_f128* outer_rounds,
                         bitz_f128* outer_claims, bitz_f128* inner_rounds, bitz_f128* inner_point,
                         bitz_f128* matrix_eval, bitz_f128* final_claim) {
    if (!t || !h || !digest || !az || !bz || !cz || !(assignment || bits) || !outer_claims || !matrix_eval ||
        !final_claim ||
        (row_vars && !outer_rounds) || (col_vars && (!inner_rounds || !inner_point)))
        return fail(BITZ_ERR_ARG, "null pointer");
    // piop.rs:74-83: dimensions before any transcript call.
    if ((int)row_vars != h->row_vars) return fail(BITZ_ERR_PRODUCT_DIMS, "products do not span the row domain");
    if ((int)col_vars != h->col_vars) return fail(BITZ_ERR_ASSIGNMENT_DIMS, "assignment does not span the columns");
    hipStream_t s = h->so.s;
    const bool dev = (flags & BITZ_DEVICE) != 0;
    // piop.rs:85-90: digest, tau, make_equality_factors (crates/poly/src/eq.rs:107-115).
    t->t.absorb(0, digest, 32);
    std::vector<f128> tau(row_vars);
    for (uint32_t i = 0; i < row_vars; ++i) tau[i] = t->t.squeeze();
    const uint32_t split = row_vars / 2;
    auto host_eq = [](const f128* r, uint32_t k) {  // eq_table (eq.rs:63-99)
        std::vector<f128> tbl((size_t)1 << k, f128{0, 0});
        tbl[0] = hostf::one();
        for (uint32_t i = 0; i < k; ++i) {
            const size_t half = (size_t)1 << i;
            for (size_t j = 0; j < half; ++j) {
                const f128 one = hostf::mul(tbl[j], r[i]);
                tbl[j] = hostf::add(tbl[j], one);
                tbl[j + half] = one;
            }
        }
        return tbl;
    };
    const std::vector<f128> lo = host_eq(tau.data(), split), hi = host_eq(tau.data() + split, row_vars - split);
    Dev dlo, dhi, drx, dD, dh;
    const bitz_f128* eq_lo = reinterpret_cast<const bitz_f128*>(lo.data());
    const bitz_f128* eq_hi = reinterpret_cast<const bitz_f128*>(hi.data());
    if (dev) {
        TRY(upload(dlo, eq_lo, lo.size(), 0, s));
        TRY(upload(dhi, eq_hi, hi.size(), 0, s));
        HIP_TRY(hipStreamSynchronize(s));
        eq_lo = reinterpret_cast<const bitz_f128*>(dlo.p);
        eq_hi = reinterpret_cast<const bitz_f128*>(dhi.p);
    }
    // piop.rs:91: the outer sumcheck from claim 0.
    std::vector<bitz_f128> rx(std::max<uint32_t>(row_vars, 1));
    bitz_f128 outer_final;
    TRY(bitz_prove_outer_sumcheck(t, bitz_f128{0, 0}, eq_lo, split, eq_hi, row_vars - split, az, row_vars, bz,
                                  row_vars, cz, row_vars, flags & BITZ_DEVICE, outer_rounds, rx.data(), &outer_final,
                                  outer_claims));
    // piop.rs:94-98: rho, the inner claim, D = bind_and_batch(r_x, rho), on the device.
    const f128 rho = t->t.squeeze();
    const f128 a = F(outer_claims[0]), b = F(outer_claims[1]), c = F(outer_claims[2]);
    const f128 inner_claim = hostf::add(a, hostf::add(hostf::mul(rho, b), hostf::mul(hostf::mul(rho, rho), c)));
    TRY(upload(drx, rx.data(), std::max<uint32_t>(row_vars, 1), 0, s));
    TRY(dD.alloc(h->cols * sizeof(f128)));
    HIP_TRY(hipStreamSynchronize(s));
    TRY(bitz_bind_and_batch(h, reinterpret_cast<const bitz_f128*>(drx.p), row_vars, B(rho),
                            reinterpret_cast<bitz_f128*>(dD.p), BITZ_DEVICE));
    // piop.rs:99: the inner sumcheck against the assignment.
    bitz_f128 evals[2];
    if (bits) {
        // The bits' own placement (flags) and D on the device.
        bitz_inner* ih = nullptr;
        Dev dbits;
        const uint64_t* bv = bits;
        if (!dev) {
            const size_t words = (h->cols + 63) / 64;
            TRY(dbits.alloc(words * 8));
            HIP_TRY(hipMemcpyAsync(dbits.p, bits, words * 8, hipMemcpyHostToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));
            bv = static_cast<const uint64_t*>(dbits.p);
        }
        TRY(bitz_inner_begin_packed(&ih, reinterpret_cast<const bitz_f128*>(dD.p), col_vars, bv, col_vars,
                                    BITZ_DEVICE));
        TRY(prove_inner_rounds(ih, t, B(inner_claim), col_vars, inner_rounds, inner_point, final_claim, evals));
    } else {
        const bitz_f128* hv = assignment;
        if (!dev) {
            TRY(upload(dh, assignment, h->cols, 0, s));
            HIP_TRY(hipStreamSynchronize(s));
            hv = reinterpret_cast<const bitz_f128*>(dh.p);
        }
        TRY(prove_inner_impl(t, B(inner_claim), reinterpret_cast<const bitz_f128*>(dD.p), col_vars, hv, col_vars,
                             BITZ_DEVICE, inner_rounds, inner_point, final_claim, evals));
    }
    *matrix_eval = evals[0];  // ScaledMleEvaluationClaim (piop.rs:101-105): D(r_y) scales h(r_y)
    return BITZ_OK;
}

extern "C" int32_t bitz_prove_spartan_piop(bitz_transcript* t, bitz_matrices* h, const uint8_t digest[32],
                                           const bitz_f128* az, const bitz_f128* bz, const bitz_f128* cz,
                                           uint32_t row_vars, const bitz_f128* assignment, uint32_t col_vars,
                                           uint32_t flags, bitz_f128* outer_rounds, bitz_f128* outer_claims,
                                           bitz_f128* inner_rounds, bitz_f128* inner_point, bitz_f128* matrix_eval,
                                           bitz_f128* final_claim) {
    if (!assignment) return fail(BITZ_ERR_ARG, "null pointer");
    return piop_impl(t, h, digest, az, bz, cz, row_vars, assignment, nullptr, col_vars, flags, outer_rounds,
                     outer_claims, inner_rounds, inner_point, matrix_eval, final_claim);
}

extern "C" int32_t bitz_prove_spartan_piop_packed(bitz_transcript* t, bitz_matrices* h, const uint8_t digest[32],
                                                  const bitz_f128* az, const bitz_f128* bz, const bitz_f128* cz,
                                                  uint32_t row_vars, const uint64_t* assignment_bits,
                                                  uint32_t col_vars, uint32_t flags, bitz_f128* outer_rounds,
                                                  bitz_f128* outer_claims, bitz_f128* inner_rounds,
                                                  bitz_f128* inner_point, bitz_f128* matrix_eval,
                                                  bitz_f128* final_claim) {
    if (!assignment_bits) return fail(BITZ_ERR_ARG, "null pointer");
    return piop_impl(t, h, digest, az, bz, cz, row_vars, nullptr, assignment_bits, col_vars, flags, outer_rounds,
                     outer_claims, inner_rounds, inner_point, matrix_eval, final_claim);
}

extern "C" void bitz_matrices_free(bitz_matrices* h) { delete h; }

// Whether provers with a Keccak transcript draw its challenges on the
// device (bitz_set_device_fiat_shamir); on by default.
static std::atomic<bool> g_device_fs{true};

// ============================================================ PIOP engine ==
// Many prove_spartan_piop proofs advanced together, one round each per step.
// Every proof keeps its own transcript, tables and state; a step gathers the
// round each proof is at, launches every round of a kind in one kernel
// (outer round 0, outer fold rounds, inner round 0, inner fold rounds; the
// final folds and bind_and_batch per proof), reads all sums back at once,
// then each proof's transcript absorbs its round polynomial and squeezes its
// next challenge on the host.  A proof submitted between steps starts at the
// next step.  The transcript order is bitz_prove_spartan_piop's.
namespace {

// kDevDone: a device-transcript proof whose last action is in the step in
// flight (its outputs come back with that step).
enum class JobPhase { kOuter, kOuterFinal, kBind, kInner, kInnerFinal, kDevDone, kDone };

// Positions per step: a device-transcript proof advances by up to this many
// actions (rounds) per step; a host-transcript proof by one.  Measured on the
// serving mix: 2, 4, 6, 8, 12, 16 positions gave 466, 455, 463, 430, 403, 389
// K blocks/s, with p50 at 1000 requests/s of 1.35, 1.21, 1.16, 1.16, 1.15,
// 1.13 ms (shorter steps let proofs join and leave sooner; longer ones
// spread each step's host work over more rounds).
// Positions (rounds per proof) a step may advance, at most
// kEngineMaxPositions.  By default (g_engine_positions = 0) a step with few
// proofs takes kEnginePositions, so that a proof needs fewer steps (latency at
// low load), and one with kEngineBusyProofs or more takes kEngineBusyPositions,
// so that proofs join and leave sooner (throughput under load: 485-501 K
// blocks/s at 4 positions against 467-478 K at 6, while p50 at 1000
// requests/s was 1.26 ms at 4 against 1.17 at 6); bitz_set_engine_positions
// fixes it.
constexpr int kEnginePositions = 6;
constexpr int kEngineBusyPositions = 4;
constexpr size_t kEngineBusyProofs = 8;
constexpr int kEngineMaxPositions = 16;
static std::atomic<int> g_engine_positions{0};

// Words of a proof's slot in a step's read-back (a bit round's take
// kInnerBitsAccWords): round sums, round 0's live length at words 8-9.
constexpr size_t kEngineSlotWords = 16;

struct PiopJob {
    uint64_t tag = 0;
    bitz_transcript* t = nullptr;
    bitz_matrices* m = nullptr;
    const f128* hv = nullptr;  // the assignment (device)
    // Or the assignment as packed bits (device): inner rounds 0 .. kb run the
    // bit kernels, round kb writes h_kb as elements (see bitz_inner).
    const uint64_t* hbits = nullptr;
    int kb = -1;
    int n = 0, nc = 0;
    JobPhase phase = JobPhase::kOuter;
    int round = 0;  // next round to launch in the current sumcheck
    bool q1 = true;
    f128 claim{0, 0}, r{0, 0}, rho{0, 0};
    std::vector<f128> lo, hi;
    Dev dlo, dhi, dD, iP, iQ;
    bitz_outer* outer = nullptr;
    const f128* icur[2] = {nullptr, nullptr};
    int iwhere = 0;
    size_t ilen = 0, icapP = 0, icapQ = 0;
    size_t ilive = 1;  // entries of the inner tables that may be nonzero (round 0 measures it)
    f128* pend[3] = {nullptr, nullptr, nullptr};  // outputs of the fold launched this step
    std::vector<bitz_f128> outer_rounds, rx, inner_rounds, inner_point;
    bitz_f128 outer_claims[3] = {}, matrix_eval{}, final_claim{};
    int32_t status = BITZ_OK;
    // Device-transcript proofs (a packed assignment and a Keccak transcript
    // with device Fiat-Shamir on): one device block holds the running state
    // (bitz::PiopFsState), the outputs, every round's sums and r's table.
    bool devfs = false;
    Dev fsb;
    enum Out { kOutOuter, kOutRx, kOutClaims, kOutInner, kOutPoint, kOutMatrixEval, kOutFinal, kOutCount };
    size_t out_off[kOutCount] = {};  // element offsets of the outputs, after the state
    size_t back_bytes = 0;           // the state and the outputs: what comes back at the end
    size_t so_off = 0, sb_off = 0, sg_off = 0, sf_off = 0, rt_off = 0;  // byte offsets in fsb
    // Pinned staging of the proof's uploads (kept until the proof is gone, so
    // the copies are asynchronous): the packed bits; the eq factors and the
    // state's first image.
    PinnedBuf pin_bits, pin_setup;
    // Set-up work (uploads, products, tables) runs on the proof's own stream
    // from the submitting thread; the engine's stream waits for `ready`
    // before the proof's first kernels.
    StreamOwner prep;
    hipEvent_t ready = nullptr;
    Dev own_bits, own_prod;  // bitz_piop_engine_submit_assignment: the packed bits and Az, Bz, Cz
    char* fs_base() const { return static_cast<char*>(fsb.p); }
    bitz::PiopFsState* fs_state() const { return reinterpret_cast<bitz::PiopFsState*>(fs_base()); }
    f128* fs_out(int w) const {
        return reinterpret_cast<f128*>(fs_base() + sizeof(bitz::PiopFsState)) + out_off[w];
    }
    uint32_t* fs_slot_outer(int k) const {  // k = n: the final A(r), B(r), C(r)
        return reinterpret_cast<uint32_t*>(fs_base() + so_off) + 12 * (size_t)k;
    }
    uint32_t* fs_slot_bits(int K) const {
        return reinterpret_cast<uint32_t*>(fs_base() + sb_off) + bitz::kInnerBitsAccWords * (size_t)K;
    }
    uint32_t* fs_slot_gen(int k) const { return reinterpret_cast<uint32_t*>(fs_base() + sg_off) + 8 * (size_t)k; }
    uint32_t* fs_slot_final() const { return reinterpret_cast<uint32_t*>(fs_base() + sf_off); }
    uint4* fs_rtab() const { return reinterpret_cast<uint4*>(fs_base() + rt_off); }
    size_t fs_back_bytes() const { return back_bytes; }
    ~PiopJob() {
        bitz_outer_free(outer);
        if (ready) (void)hipEventDestroy(ready);
    }
};

}  // namespace

// A step in flight: its proofs, where each host-transcript proof's sums and
// each finishing proof's outputs land in the pinned buffer `buf`, and the
// event recorded after its last operation.
struct EngineStep {
    std::vector<PiopJob*> act;
    std::vector<size_t> slot_off;  // word offset of each proof's slot (and the total)
    std::vector<std::pair<PiopJob*, size_t>> fin_back;
    int buf = 0;
};

struct bitz_piop_engine {
    std::vector<std::unique_ptr<PiopJob>> jobs;
    // Steps in flight, oldest first (the worker keeps one; the two staging
    // buffers let a caller of bitz_piop_engine_launch alternate them).
    std::deque<EngineStep> steps;
    Dev slots, stage[2], rt4;
    // Bind buffers per circuit, on this engine's stream: binds on different
    // engines never wait for one another.
    std::map<const bitz_matrices*, std::unique_ptr<BindScratch>> binds;
    PinnedBuf pin[2];  // per step buffer: the uploads, then what comes back (pooled across engines)
    hipEvent_t step_done[2] = {nullptr, nullptr};
    int next_buf = 0;
    // Background stepping (bitz_piop_engine_run): a thread of the engine's
    // own steps while proofs are in flight; submit hands it proofs through
    // `incoming`, finished ones wait in `finished` for take.
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<PiopJob>> incoming, finished;
    std::thread worker;
    bool running = false, stop = false;
    int32_t worker_rc = BITZ_OK;
    std::string worker_err;
    // HIP events around the step's batched inner fold and around the whole
    // step's device work (when profiling), read once the step has synchronised.
    Timed t_inner_fold[kEngineMaxPositions], t_step, t_bits[kEngineMaxPositions];
    StreamOwner so;
    ~bitz_piop_engine() {
        if (running) {
            {
                std::lock_guard<std::mutex> lock(mu);
                stop = true;
            }
            cv.notify_all();
            worker.join();
        }
        (void)hipStreamSynchronize(so.s);
        for (int b = 0; b < 2; ++b)
            if (step_done[b]) (void)hipEventDestroy(step_done[b]);
    }
};

extern "C" int32_t bitz_piop_engine_create(bitz_piop_engine** out) {
    if (!out) return fail(BITZ_ERR_ARG, "null pointer");
    auto* e = new bitz_piop_engine;
    if (int32_t rc = e->so.create()) {
        delete e;
        return rc;
    }
    *out = e;
    return BITZ_OK;
}

extern "C" void bitz_piop_engine_free(bitz_piop_engine* e) { delete e; }

// A device-transcript proof's buffers, all allocated up front (the host
// never waits for a round): the inner tables, and one block with the state,
// the outputs, a slot per round (zeroed here) and r's table.
static int32_t devfs_setup(PiopJob* j, hipStream_t s) {
    const int n = j->n, nc = j->nc;
    const size_t N = (size_t)1 << nc;
    j->ilen = N;
    j->ilive = N;
    j->icapP = std::max<size_t>(N / 2, 1);
    j->icapQ = std::max<size_t>(N / 4, 1);
    TRY(j->dD.alloc(N * sizeof(f128)));
    TRY(j->iP.alloc(2 * j->icapP * sizeof(f128)));
    TRY(j->iQ.alloc(2 * j->icapQ * sizeof(f128)));
    const size_t counts[PiopJob::kOutCount] = {4 * (size_t)n, (size_t)n, 3, 3 * (size_t)nc, (size_t)nc, 1, 1};
    size_t o = 0;
    for (int w = 0; w < PiopJob::kOutCount; ++w) {
        j->out_off[w] = o;
        o += counts[w];
    }
    j->back_bytes = sizeof(bitz::PiopFsState) + o * sizeof(f128);
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    j->so_off = up(j->back_bytes);
    j->sb_off = j->so_off + up(12 * 4 * (size_t)(n + 1));
    j->sg_off = j->sb_off + bitz::kInnerBitsAccWords * 4 * (size_t)(j->kb + 1);
    j->sf_off = j->sg_off + up(8 * 4 * (size_t)nc);
    j->rt_off = j->sf_off + 256;
    const size_t total = j->rt_off + kConstTable4Entries * sizeof(uint4);
    TRY(j->fsb.alloc(total));
    HIP_TRY(hipMemsetAsync(j->fs_base() + j->so_off, 0, j->rt_off - j->so_off, s));
    // The state's first image, staged after the eq factors (engine_submit).
    char* img = j->pin_setup.c() + (j->lo.size() + j->hi.size()) * sizeof(f128);
    memset(img, 0, sizeof(bitz::PiopFsState));
    auto* st = reinterpret_cast<bitz::PiopFsState*>(img);
    st->ts = j->t->t.keccak;  // the duplex after the digest and tau
    st->S = hostf::one();
    st->mono[0] = hostf::one();
    HIP_TRY(hipMemcpyAsync(j->fsb.p, img, sizeof(bitz::PiopFsState), hipMemcpyHostToDevice, s));
    j->devfs = true;
    return BITZ_OK;
}

static int32_t engine_submit(bitz_piop_engine* e, uint64_t tag, bitz_transcript* t, bitz_matrices* m,
                             const uint8_t digest[32], const bitz_f128* az, const bitz_f128* bz, const bitz_f128* cz,
                             uint32_t row_vars, const bitz_f128* assignment, const uint64_t* bits,
                             uint32_t col_vars, std::unique_ptr<PiopJob> job = nullptr) {
    if (!e || !t || !m || !digest || !az || !bz || !cz || !(assignment || bits))
        return fail(BITZ_ERR_ARG, "null pointer");
    // piop.rs:74-83: dimensions before any transcript call.
    if ((int)row_vars != m->row_vars) return fail(BITZ_ERR_PRODUCT_DIMS, "products do not span the row domain");
    if ((int)col_vars != m->col_vars) return fail(BITZ_ERR_ASSIGNMENT_DIMS, "assignment does not span the columns");
    if (row_vars == 0 || col_vars == 0 || row_vars > 32)
        return fail(BITZ_ERR_ARG, "the engine proves 1 <= row_vars <= 32 and col_vars >= 1");
    {
        std::lock_guard<std::mutex> lock(e->mu);
        for (const auto* list : {&e->jobs, &e->incoming, &e->finished})
            for (const auto& j : *list)
                if (j->tag == tag) return fail(BITZ_ERR_ARG, "tag in use");
    }
    std::unique_ptr<PiopJob> j = job ? std::move(job) : std::make_unique<PiopJob>();
    if (!j->prep.s) TRY(j->prep.create());
    j->tag = tag;
    j->t = t;
    j->m = m;
    j->hv = reinterpret_cast<const f128*>(assignment);
    j->n = (int)row_vars;
    j->nc = (int)col_vars;
    if (bits) {
        j->hbits = bits;
        j->kb = std::min<int>(bitz::kInnerBitsMaxRound, (int)col_vars - 1);
    }
    hipStream_t s = j->prep.s;  // set-up on the proof's own stream
    // piop.rs:85-90: digest, tau, make_equality_factors (crates/poly/src/eq.rs:107-115).
    t->t.absorb(0, digest, 32);
    std::vector<f128> tau(row_vars);
    for (uint32_t i = 0; i < row_vars; ++i) tau[i] = t->t.squeeze();
    const uint32_t split = row_vars / 2;
    auto host_eq = [](const f128* r, uint32_t k) {  // eq_table (eq.rs:63-99)
        std::vector<f128> tbl((size_t)1 << k, f128{0, 0});
        tbl[0] = hostf::one();
        for (uint32_t i = 0; i < k; ++i) {
            const size_t half = (size_t)1 << i;
            for (size_t x = 0; x < half; ++x) {
                const f128 one = hostf::mul(tbl[x], r[i]);
                tbl[x] = hostf::add(tbl[x], one);
                tbl[x + half] = one;
            }
        }
        return tbl;
    };
    j->lo = host_eq(tau.data(), split);
    j->hi = host_eq(tau.data() + split, row_vars - split);
    {
        // Through pinned staging: the copies stay asynchronous.
        const size_t lb = j->lo.size() * sizeof(f128), hb = j->hi.size() * sizeof(f128);
        TRY(j->pin_setup.alloc(lb + hb + sizeof(bitz::PiopFsState)));
        memcpy(j->pin_setup.c(), j->lo.data(), lb);
        memcpy(j->pin_setup.c() + lb, j->hi.data(), hb);
        TRY(j->dlo.alloc(lb));
        TRY(j->dhi.alloc(hb));
        HIP_TRY(hipMemcpyAsync(j->dlo.p, j->pin_setup.c(), lb, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(j->dhi.p, j->pin_setup.c() + lb, hb, hipMemcpyHostToDevice, s));
    }
    // The outer sumcheck from claim 0 over the eq-factored kernels; the
    // factors are eq tables of tau by construction, so no device check.
    const auto t_outer = std::chrono::steady_clock::now();
    TRY(outer_begin_impl(&j->outer, reinterpret_cast<const bitz_f128*>(j->dlo.p), split,
                         reinterpret_cast<const bitz_f128*>(j->dhi.p), row_vars - split, az, row_vars, bz, row_vars,
                         cz, row_vars, BITZ_DEVICE, false));
    bitz_outer* h = j->outer;
    if (!h->eqmode || !outer_eq_constants(h, tau))
        return fail(BITZ_ERR_ARG, "eq factors outside the eq-factored kernels' range");
    h->started = true;
    if (h->htab_total)
        hipLaunchKernelGGL(bitz::k_htabs4, dim3((unsigned)h->htab_total), dim3(1024), 0, s, h->eq_in[0], h->ell,
                           h->eq_in[1], h->hdesc, static_cast<uint4*>(h->htabs4.p));
    TRY(check_launch());
    j->outer_rounds.resize(4 * (size_t)row_vars);
    j->rx.resize(row_vars);
    j->inner_rounds.resize(3 * (size_t)col_vars);
    j->inner_point.resize(col_vars);
    prof_add_host("submit_outer", t_outer);
    const auto t_devfs = std::chrono::steady_clock::now();
    if (bits && t->t.type == Transcript::Type::kKeccak && g_device_fs) TRY(devfs_setup(j.get(), s));
    prof_add_host("submit_devfs", t_devfs);
    HIP_TRY(hipEventCreateWithFlags(&j->ready, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(j->ready, s));
    if (e->running) {
        {
            std::lock_guard<std::mutex> lock(e->mu);
            e->incoming.push_back(std::move(j));
        }
        e->cv.notify_all();
    } else {
        e->jobs.push_back(std::move(j));
    }
    return BITZ_OK;
}

extern "C" int32_t bitz_piop_engine_submit(bitz_piop_engine* e, uint64_t tag, bitz_transcript* t, bitz_matrices* m,
                                           const uint8_t digest[32], const bitz_f128* az, const bitz_f128* bz,
                                           const bitz_f128* cz, uint32_t row_vars, const bitz_f128* assignment,
                                           uint32_t col_vars) {
    if (!assignment) return fail(BITZ_ERR_ARG, "null pointer");
    return engine_submit(e, tag, t, m, digest, az, bz, cz, row_vars, assignment, nullptr, col_vars);
}

extern "C" int32_t bitz_piop_engine_submit_packed(bitz_piop_engine* e, uint64_t tag, bitz_transcript* t,
                                                  bitz_matrices* m, const uint8_t digest[32], const bitz_f128* az,
                                                  const bitz_f128* bz, const bitz_f128* cz, uint32_t row_vars,
                                                  const uint64_t* assignment_bits, uint32_t col_vars) {
    if (!assignment_bits) return fail(BITZ_ERR_ARG, "null pointer");
    return engine_submit(e, tag, t, m, digest, az, bz, cz, row_vars, nullptr, assignment_bits, col_vars);
}

extern "C" int32_t bitz_piop_engine_submit_assignment(bitz_piop_engine* e, uint64_t tag, bitz_transcript* t,
                                                      bitz_matrices* m, const uint8_t digest[32],
                                                      const uint64_t* words, size_t bit_len, size_t columns) {
    if (!e || !t || !m || !digest || (bit_len && !words)) return fail(BITZ_ERR_ARG, "null pointer");
    // build_assignment_mle (matrix.rs:392-420) checks, then the column domain.
    if (bit_len != columns) return fail(BITZ_ERR_ASSIGNMENT_LEN, "assignment length differs from the columns");
    if (columns == 0 || !(words[0] & 1)) return fail(BITZ_ERR_ASSIGNMENT_CONST, "the first assignment bit must be one");
    size_t padded = 1;
    while (padded < columns) padded <<= 1;
    if (padded != m->cols) return fail(BITZ_ERR_ASSIGNMENT_DIMS, "assignment does not span the columns");
    const auto t_sub = std::chrono::steady_clock::now();
    auto j = std::make_unique<PiopJob>();
    TRY(j->prep.create());
    hipStream_t s = j->prep.s;
    // The packed bits (zero past the assignment) and Az, Bz, Cz from them.
    const size_t nw = (padded + 63) / 64, used = (bit_len + 63) / 64;
    TRY(j->pin_bits.alloc(nw * 8));
    uint64_t* packed = reinterpret_cast<uint64_t*>(j->pin_bits.p);
    memcpy(packed, words, used * 8);
    if (bit_len % 64) packed[used - 1] &= (~0ull) >> (64 - bit_len % 64);
    memset(packed + used, 0, (nw - used) * 8);
    TRY(j->own_bits.alloc(nw * 8));
    const auto t_up = std::chrono::steady_clock::now();
    HIP_TRY(hipMemcpyAsync(j->own_bits.p, packed, nw * 8, hipMemcpyHostToDevice, s));
    prof_add_host("submit_upload", t_up);
    const uint64_t* bits = static_cast<const uint64_t*>(j->own_bits.p);
    TRY(j->own_prod.alloc(3 * m->rows_padded * sizeof(f128)));
    for (int k = 0; k < 3; ++k)
        hipLaunchKernelGGL(bitz::k_r1cs_rows_bits, dim3(grid_for(m->rows_padded)), dim3(kBlock), 0, s, m->rs[k].u32(),
                           m->cs[k].u32(), m->vs[k].f(), bits, m->rows, j->own_prod.f() + k * m->rows_padded,
                           m->rows_padded);
    TRY(check_launch());
    prof_add_host("submit_prepare", t_sub);
    const bitz_f128* prod = reinterpret_cast<const bitz_f128*>(j->own_prod.p);
    const int32_t rc = engine_submit(e, tag, t, m, digest, prod, prod + m->rows_padded, prod + 2 * m->rows_padded,
                                     (uint32_t)m->row_vars, nullptr, bits, (uint32_t)m->col_vars, std::move(j));
    prof_add_host("submit_total", t_sub);
    return rc;
}

// Inner round K's challenges r_0 .. r_{K-1} as elements.
static std::vector<f128> job_inner_point(const PiopJob* j, int K) {
    std::vector<f128> r((size_t)std::max(K, 1));
    for (int i = 0; i < K; ++i) r[i] = F(j->inner_point[i]);
    return r;
}

// The kernels of one position of a step: every round of one kind across the
// proofs at that position in one launch.
namespace {
struct EnginePos {
    std::vector<bitz::RoundDesc> o_r0, o_fold;
    std::vector<uint32_t> o_r0_map, o_fold_map;
    std::vector<bitz::InnerRoundDesc> i_r0, i_fold;
    std::vector<uint32_t> i_r0_map, i_fold_map;
    std::vector<bitz::InnerBitsDesc> i_bits;  // every bit round K of the position, one launch
    std::vector<uint32_t> i_bits_map;
    std::vector<bitz::PiopFsDesc> fs;
    std::vector<size_t> binds;       // proofs (indices into the step's list) that bind here
    std::vector<size_t> host_final;  // host-transcript proofs' final folds (position 0)
    // staging offsets
    size_t or_o = 0, orm_o = 0, of_o = 0, ofm_o = 0, ir_o = 0, irm_o = 0, if_o = 0, ifm_o = 0, fin_o = 0, fs_o = 0;
    size_t ib_o = 0, ibm_o = 0;
    void add_bits(int K, bitz::InnerBitsDesc d) {
        d.blocks = bits_grid(d.pairs);
        d.block_base = (uint32_t)i_bits_map.size();
        d.k = K;
        i_bits_map.insert(i_bits_map.end(), d.blocks, (uint32_t)i_bits.size());
        i_bits.push_back(d);
    }
    void add_fold(bitz::InnerRoundDesc d) {
        d.blocks = inner_grid(d.pairs);
        d.block_base = (uint32_t)i_fold_map.size();
        i_fold_map.insert(i_fold_map.end(), d.blocks, (uint32_t)i_fold.size());
        i_fold.push_back(d);
    }
    void add_outer(int k, const bitz_outer* h, bitz::RoundDesc d) {
        const auto& g = h->geo[k];
        d.a = h->cur[0];
        d.b = h->cur[1];
        d.c = h->cur[2];
        d.la = h->table_ptr(g.la.table);
        d.lb = h->table_ptr(g.lb.table);
        d.h_tabs4 = static_cast<const uint4*>(h->htabs4.p) + g.htab_off * kConstTable4Entries;
        d.threads = g.threads;
        d.iters = g.iters;
        d.la_bits = g.la_bits;
        d.la_shift = g.la.a - h->table_base(g.la.table);
        d.lb_shift = g.lb.a - h->table_base(g.lb.table);
        auto& descs = k == 0 ? o_r0 : o_fold;
        auto& map = k == 0 ? o_r0_map : o_fold_map;
        const unsigned block = k == 0 ? bitz::kR0Block : bitz::kFoldBlock;
        d.block_base = (uint32_t)map.size();
        map.insert(map.end(), (g.threads + block - 1) / block, (uint32_t)descs.size());
        descs.push_back(d);
    }
};

// A device-transcript proof's next action, appended to position P: its
// round's kernels and the step of its transcript (k_piop_fs), then its host
// cursor moves on (nothing it needs comes back from the device).
void devfs_emit(PiopJob* j, size_t i, EnginePos& P) {
    bitz::PiopFsState* st = j->fs_state();
    switch (j->phase) {
        case JobPhase::kOuter: {
            bitz_outer* h = j->outer;
            const int k = j->round;
            bitz::RoundDesc d{};
            if (k > 0) {
                Dev& dst = (h->where == 1) ? h->Q : h->P;
                const size_t cap = (h->where == 1) ? h->capQ : h->capP;
                d.ao = dst.f();
                d.bo = dst.f() + cap;
                d.co = dst.f() + 2 * cap;
                d.r_tab4 = j->fs_rtab();
            }
            const bool q1 = k == 0 || hostf::is_zero(h->tau[k]);
            d.q1 = q1 ? 1 : 0;
            d.acc = j->fs_slot_outer(k);
            P.add_outer(k, h, d);
            bitz::PiopFsDesc f{};
            f.st = st;
            f.kind = bitz::kFsOuterRound;
            f.k = k;
            f.K = -1;
            f.have_q1 = q1 ? 1 : 0;
            f.sums = j->fs_slot_outer(k);
            f.round_out = j->fs_out(PiopJob::kOutOuter) + 4 * (size_t)k;
            f.point_out = j->fs_out(PiopJob::kOutRx) + k;
            f.rtab = j->fs_rtab();
            f.tau = h->tau[k];
            f.tau_inv = h->tau_inv[k];
            f.scale = h->scale_big[k];
            {
                f128 one_tau = h->tau[k];
                one_tau.lo ^= 1u;
                f.scale_t = hostf::mul(f.scale, one_tau);
                f.scale_ti = q1 ? f128{0, 0} : hostf::mul(f.scale_t, h->tau_inv[k]);
            }
            P.fs.push_back(f);
            if (k > 0) {
                h->cur[0] = d.ao;
                h->cur[1] = d.bo;
                h->cur[2] = d.co;
                h->where = (h->where == 1) ? 2 : 1;
                h->len /= 2;
                h->round += 1;
            }
            j->round = k + 1;
            if (j->round == j->n) j->phase = JobPhase::kOuterFinal;
            break;
        }
        case JobPhase::kOuterFinal: {
            bitz_outer* h = j->outer;
            bitz::PiopFsDesc f{};
            for (int u = 0; u < 3; ++u) f.fin_tab[u] = h->cur[u];
            f.fin_r = j->fs_out(PiopJob::kOutRx) + (j->n - 1);
            f.fin_count = 3;
            f.st = st;
            f.kind = bitz::kFsOuterFinal;
            f.sums = j->fs_slot_outer(j->n);
            f.round_out = j->fs_out(PiopJob::kOutClaims);
            P.fs.push_back(f);
            j->phase = JobPhase::kBind;
            j->round = 0;
            break;
        }
        case JobPhase::kBind: {
            // bind_and_batch at the outer point and rho (device), then inner round 0.
            P.binds.push_back(i);
            bitz::InnerBitsDesc d{};
            d.d = j->dD.f();
            d.bits = j->hbits;
            d.pairs = j->ilen / 2;
            d.live = j->ilen;
            d.acc = j->fs_slot_bits(0);
            if (j->kb == 0) {  // one round: h as elements for the final fold
                d.wout = j->iP.f() + j->icapP;
                d.mono_dev = st->mono;
            }
            P.add_bits(0, d);
            bitz::PiopFsDesc f{};
            f.st = st;
            f.kind = bitz::kFsInnerRound;
            f.k = 0;
            f.K = 0;
            f.mono_next = 0 < j->kb;
            f.sums = j->fs_slot_bits(0);
            f.round_out = j->fs_out(PiopJob::kOutInner);
            f.point_out = j->fs_out(PiopJob::kOutPoint);
            f.rtab = j->fs_rtab();
            P.fs.push_back(f);
            j->icur[0] = j->dD.f();
            j->icur[1] = j->kb == 0 ? d.wout : nullptr;
            j->round = 1;
            j->phase = j->nc == 1 ? JobPhase::kInnerFinal : JobPhase::kInner;
            break;
        }
        case JobPhase::kInner: {
            const int k = j->round;
            const size_t next_len = j->ilen / 2;
            Dev& dst = (j->iwhere == 1) ? j->iQ : j->iP;
            const size_t cap = (j->iwhere == 1) ? j->icapQ : j->icapP;
            bitz::PiopFsDesc f{};
            f.st = st;
            f.kind = bitz::kFsInnerRound;
            f.k = k;
            f.round_out = j->fs_out(PiopJob::kOutInner) + 3 * (size_t)k;
            f.point_out = j->fs_out(PiopJob::kOutPoint) + k;
            f.rtab = j->fs_rtab();
            f128* wout = nullptr;
            if (k <= j->kb) {
                bitz::InnerBitsDesc d{};
                d.d = j->icur[0];
                d.bits = j->hbits;
                d.dout = dst.f();
                wout = k == j->kb ? dst.f() + cap : nullptr;
                d.wout = wout;
                d.mono_dev = st->mono;
                d.pairs = next_len / 2;  // bounds; the kernel takes the live length from the device
                d.live = j->ilen;
                d.live0 = &st->live;
                d.rtab4 = j->fs_rtab();
                d.acc = j->fs_slot_bits(k);
                P.add_bits(k, d);
                f.K = k;
                f.mono_next = k < j->kb;
                f.sums = j->fs_slot_bits(k);
            } else {
                bitz::InnerRoundDesc d{};
                d.d = j->icur[0];
                d.w = j->icur[1];
                d.dout = dst.f();
                d.wout = dst.f() + cap;
                wout = d.wout;
                d.pairs = next_len / 2;
                d.live = j->ilen;
                d.live0 = &st->live;
                d.live_shift = k - 1;
                d.rtab4 = j->fs_rtab();
                d.acc = j->fs_slot_gen(k);
                P.add_fold(d);
                f.K = -1;
                f.sums = j->fs_slot_gen(k);
            }
            P.fs.push_back(f);
            j->icur[0] = dst.f();
            j->icur[1] = wout;
            j->iwhere = (j->iwhere == 1) ? 2 : 1;
            j->ilen = next_len;
            j->round = k + 1;
            if (j->round == j->nc) j->phase = JobPhase::kInnerFinal;
            break;
        }
        case JobPhase::kInnerFinal: {
            bitz::PiopFsDesc f{};
            f.fin_tab[0] = j->icur[0];
            f.fin_tab[1] = j->icur[1];
            f.fin_r = j->fs_out(PiopJob::kOutPoint) + (j->nc - 1);
            f.fin_count = 2;
            f.st = st;
            f.kind = bitz::kFsInnerFinal;
            f.sums = j->fs_slot_final();
            f.round_out = j->fs_out(PiopJob::kOutMatrixEval);
            f.final_out = j->fs_out(PiopJob::kOutFinal);
            P.fs.push_back(f);
            j->phase = JobPhase::kDevDone;
            break;
        }
        default:
            break;
    }
}
}  // namespace

// Queues a step (every active proof) on the engine's stream behind any step
// in flight; *launched = the proofs in it (none queued when zero).
static int32_t engine_launch_step(bitz_piop_engine* e, size_t* launched) {
    const auto host_t0 = std::chrono::steady_clock::now();
    hipStream_t s = e->so.s;
    std::vector<PiopJob*> act;
    bool any_devfs = false;
    for (auto& j : e->jobs)
        if (j->phase != JobPhase::kDone && j->phase != JobPhase::kDevDone) {
            act.push_back(j.get());
            any_devfs |= j->devfs;
        }
    if (launched) *launched = act.size();
    if (act.empty()) return BITZ_OK;
    EngineStep step;
    step.buf = e->next_buf;
    // Positions: a host-transcript proof does one round per step (its
    // challenge comes from the host after the read-back); a device-transcript
    // proof does up to kEnginePositions.
    const int fixed = g_engine_positions.load();
    const int npos = !any_devfs ? 1 : fixed ? fixed
                   : act.size() >= kEngineBusyProofs ? kEngineBusyPositions : kEnginePositions;
    std::vector<EnginePos> pos(npos);
    // Each host-transcript proof's slot for the step's sums: a bit round's
    // copies (kInnerBitsAccWords), else kEngineSlotWords words.
    std::vector<size_t>& slot_off = step.slot_off;
    slot_off.assign(act.size() + 1, 0);
    for (size_t i = 0; i < act.size(); ++i) {
        const PiopJob* j = act[i];
        const bool bits = !j->devfs && j->hbits &&
                          ((j->phase == JobPhase::kBind) || (j->phase == JobPhase::kInner && j->round <= j->kb));
        slot_off[i + 1] = slot_off[i] + (j->devfs ? 0 : bits ? bitz::kInnerBitsAccWords : kEngineSlotWords);
    }
    const size_t slot_bytes = std::max<size_t>(slot_off[act.size()] * 4, 256);
    if (slot_off[act.size()] && e->slots.bytes < slot_bytes) TRY(e->slots.alloc(std::max(slot_bytes, 2 * e->slots.bytes)));
    auto slot = [&](size_t i) { return e->slots.u32() + slot_off[i]; };
    // Challenge tables to build: the r each host-transcript fold round folds by.
    std::vector<f128> consts;
    std::vector<int> table_of(act.size(), -1);
    for (size_t i = 0; i < act.size(); ++i) {
        const PiopJob* j = act[i];
        if (j->devfs) continue;
        const bool outer_fold = j->phase == JobPhase::kOuter && j->round > 0;
        const bool inner_fold = j->phase == JobPhase::kInner && j->round > 0;
        if (outer_fold || inner_fold) {
            table_of[i] = (int)consts.size();
            consts.push_back(j->r);
        }
    }
    if (e->rt4.bytes < std::max<size_t>(consts.size(), 1) * kConstTable4Entries * sizeof(uint4))
        TRY(e->rt4.alloc(std::max(std::max<size_t>(consts.size(), 1) * kConstTable4Entries * sizeof(uint4),
                                  2 * e->rt4.bytes)));
    const uint4* rt4 = static_cast<const uint4*>(e->rt4.p);
    // Host-transcript proofs: position 0.
    EnginePos& P0 = pos[0];
    for (size_t i = 0; i < act.size(); ++i) {
        PiopJob* j = act[i];
        if (j->devfs) continue;
        if (j->phase == JobPhase::kOuter) {
            bitz_outer* h = j->outer;
            const int k = j->round;
            bitz::RoundDesc d{};
            if (k > 0) {
                outer_eq_advance(h, k - 1, j->r);
                Dev& dst = (h->where == 1) ? h->Q : h->P;
                const size_t cap = (h->where == 1) ? h->capQ : h->capP;
                j->pend[0] = dst.f();
                j->pend[1] = dst.f() + cap;
                j->pend[2] = dst.f() + 2 * cap;
                d.ao = j->pend[0];
                d.bo = j->pend[1];
                d.co = j->pend[2];
                d.r_tab4 = rt4 + (size_t)table_of[i] * kConstTable4Entries;
            }
            j->q1 = k == 0 || hostf::is_zero(h->tau[k]);
            d.acc = slot(i);
            d.q1 = j->q1 ? 1 : 0;
            P0.add_outer(k, h, d);
        } else if (j->phase == JobPhase::kBind && j->hbits) {
            P0.binds.push_back(i);
            bitz::InnerBitsDesc d{};
            d.d = j->dD.f();
            d.bits = j->hbits;
            d.pairs = j->ilen / 2;
            d.live = j->ilen;
            d.acc = slot(i);
            j->pend[0] = nullptr;
            j->pend[1] = nullptr;
            if (j->kb == 0) {  // one round: h as elements for the final fold
                d.wout = j->iP.f() + j->icapP;
                d.mono[0] = hostf::one();
                j->pend[1] = d.wout;
            }
            P0.add_bits(0, d);
        } else if (j->phase == JobPhase::kInner && j->hbits && j->round <= j->kb) {
            const int K = j->round;
            const size_t next_len = j->ilen / 2;
            Dev& dst = (j->iwhere == 1) ? j->iQ : j->iP;
            const size_t cap = (j->iwhere == 1) ? j->icapQ : j->icapP;
            j->pend[0] = dst.f();
            j->pend[1] = K == j->kb ? dst.f() + cap : nullptr;
            bitz::InnerBitsDesc d{};
            d.d = j->icur[0];
            d.bits = j->hbits;
            d.dout = j->pend[0];
            d.wout = j->pend[1];
            d.pairs = std::min(next_len / 2, ((j->ilive + 1) / 2 + 1) / 2);  // the folded live prefix, in pairs
            d.live = j->ilive;
            d.rtab4 = rt4 + (size_t)table_of[i] * kConstTable4Entries;
            d.acc = slot(i);
            if (d.wout) {
                const std::vector<f128> r = job_inner_point(j, K);
                monomials(r.data(), K, d.mono);
            }
            P0.add_bits(K, d);
        } else if (j->phase == JobPhase::kBind) {
            P0.binds.push_back(i);
            bitz::InnerRoundDesc d{};
            d.d = j->dD.f();
            d.w = j->hv;
            d.pairs = j->ilen / 2;
            d.live = j->ilen;
            d.live_out = reinterpret_cast<unsigned long long*>(slot(i) + 8);  // the tables' live length
            d.acc = slot(i);
            d.blocks = inner_grid(d.pairs);
            d.block_base = (uint32_t)P0.i_r0_map.size();
            P0.i_r0_map.insert(P0.i_r0_map.end(), d.blocks, (uint32_t)P0.i_r0.size());
            P0.i_r0.push_back(d);
        } else if (j->phase == JobPhase::kInner) {
            const size_t next_len = j->ilen / 2;
            Dev& dst = (j->iwhere == 1) ? j->iQ : j->iP;
            const size_t cap = (j->iwhere == 1) ? j->icapQ : j->icapP;
            j->pend[0] = dst.f();
            j->pend[1] = dst.f() + cap;
            bitz::InnerRoundDesc d{};
            d.d = j->icur[0];
            d.w = j->icur[1];
            d.dout = j->pend[0];
            d.wout = j->pend[1];
            d.pairs = std::min(next_len / 2, ((j->ilive + 1) / 2 + 1) / 2);  // the folded live prefix, in pairs
            d.live = j->ilive;
            d.rtab4 = rt4 + (size_t)table_of[i] * kConstTable4Entries;
            d.acc = slot(i);
            P0.add_fold(d);
        } else if (j->phase == JobPhase::kOuterFinal || j->phase == JobPhase::kInnerFinal) {
            P0.host_final.push_back(i);
        }
    }
    // Device-transcript proofs: one action per position while any is left.
    std::vector<size_t> finishing;  // device-transcript proofs that end in this step
    for (size_t i = 0; i < act.size(); ++i) {
        PiopJob* j = act[i];
        if (!j->devfs) continue;
        for (int u = 0; u < npos && j->phase != JobPhase::kDevDone; ++u) devfs_emit(j, i, pos[u]);
        if (j->phase == JobPhase::kDevDone) finishing.push_back(i);
    }
    prof_add_host("piop_engine_launch_emit", host_t0);
    const auto stage_t0 = std::chrono::steady_clock::now();
    // Staging: challenges | host-transcript row points | every position's
    // descriptors and maps.
    size_t off = 0;
    auto place = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        const size_t o = off;
        off += bytes;
        return o;
    };
    const size_t c_o = place(consts.size() * sizeof(f128));
    std::vector<size_t> rx_o(act.size(), 0);
    for (size_t b : P0.binds)
        if (!act[b]->devfs) rx_o[b] = place((size_t)act[b]->n * sizeof(f128));
    for (auto& P : pos) {
        P.or_o = place(P.o_r0.size() * sizeof(bitz::RoundDesc));
        P.orm_o = place(P.o_r0_map.size() * 4);
        P.of_o = place(P.o_fold.size() * sizeof(bitz::RoundDesc));
        P.ofm_o = place(P.o_fold_map.size() * 4);
        P.ir_o = place(P.i_r0.size() * sizeof(bitz::InnerRoundDesc));
        P.irm_o = place(P.i_r0_map.size() * 4);
        P.if_o = place(P.i_fold.size() * sizeof(bitz::InnerRoundDesc));
        P.ifm_o = place(P.i_fold_map.size() * 4);
        P.ib_o = place(P.i_bits.size() * sizeof(bitz::InnerBitsDesc));
        P.ibm_o = place(P.i_bits_map.size() * 4);
        P.fs_o = place(P.fs.size() * sizeof(bitz::PiopFsDesc));
    }
    // Read-back: the host-transcript slots, then each finishing proof's state
    // and outputs.
    size_t back = slot_off[act.size()] * 4;
    std::vector<size_t> fin_back(finishing.size());
    for (size_t f = 0; f < finishing.size(); ++f) {
        back = (back + 255) & ~(size_t)255;
        fin_back[f] = back;
        back += act[finishing[f]]->fs_back_bytes();
    }
    prof_add_host("piop_engine_launch_place", stage_t0);
    const auto pin_t0 = std::chrono::steady_clock::now();
    // One pinned buffer stages the step's uploads, then receives what comes back.
    const size_t total = std::max<size_t>(off, 256);
    const size_t need = std::max(total, back);
    const int b = step.buf;
    TRY(e->pin[b].alloc(need));  // the step before last used it, and has been finished
    prof_add_host("piop_engine_launch_pin", pin_t0);
    const auto dev_t0 = std::chrono::steady_clock::now();
    TRY(e->stage[b].grow_idle(total));  // the step that used it last has been finished
    prof_add_host("piop_engine_launch_devstage", dev_t0);
    if (!e->step_done[b]) HIP_TRY(hipEventCreateWithFlags(&e->step_done[b], hipEventDisableTiming));
    prof_add_host("piop_engine_launch_stage_alloc", stage_t0, (double)need);
    char* pin = e->pin[b].c();
    char* dst = static_cast<char*>(e->stage[b].p);
    memcpy(pin + c_o, consts.data(), consts.size() * sizeof(f128));
    for (size_t b : P0.binds)
        if (!act[b]->devfs) memcpy(pin + rx_o[b], act[b]->rx.data(), (size_t)act[b]->n * sizeof(f128));
    for (auto& P : pos) {
        memcpy(pin + P.or_o, P.o_r0.data(), P.o_r0.size() * sizeof(bitz::RoundDesc));
        memcpy(pin + P.orm_o, P.o_r0_map.data(), P.o_r0_map.size() * 4);
        memcpy(pin + P.of_o, P.o_fold.data(), P.o_fold.size() * sizeof(bitz::RoundDesc));
        memcpy(pin + P.ofm_o, P.o_fold_map.data(), P.o_fold_map.size() * 4);
        memcpy(pin + P.ir_o, P.i_r0.data(), P.i_r0.size() * sizeof(bitz::InnerRoundDesc));
        memcpy(pin + P.irm_o, P.i_r0_map.data(), P.i_r0_map.size() * 4);
        memcpy(pin + P.if_o, P.i_fold.data(), P.i_fold.size() * sizeof(bitz::InnerRoundDesc));
        memcpy(pin + P.ifm_o, P.i_fold_map.data(), P.i_fold_map.size() * 4);
        memcpy(pin + P.ib_o, P.i_bits.data(), P.i_bits.size() * sizeof(bitz::InnerBitsDesc));
        memcpy(pin + P.ibm_o, P.i_bits_map.data(), P.i_bits_map.size() * 4);
        memcpy(pin + P.fs_o, P.fs.data(), P.fs.size() * sizeof(bitz::PiopFsDesc));
    }
    prof_add_host("piop_engine_launch_stage", stage_t0, (double)off);
    const auto kern_t0 = std::chrono::steady_clock::now();
    double bind_ms = 0;
    e->t_step.begin(s, "piop_engine_step", (double)act.size());  // "bytes": proofs in the step
    for (PiopJob* j : act)
        if (j->ready) {  // a new proof: its set-up, queued on its own stream, first
            HIP_TRY(hipStreamWaitEvent(s, j->ready, 0));
            (void)hipEventDestroy(j->ready);
            j->ready = nullptr;
        }
    HIP_TRY(hipMemcpyAsync(dst, pin, off, hipMemcpyHostToDevice, s));
    if (slot_off[act.size()]) HIP_TRY(hipMemsetAsync(e->slots.p, 0, slot_off[act.size()] * 4, s));
    if (!consts.empty())
        hipLaunchKernelGGL(bitz::k_const_tables4, dim3((unsigned)consts.size()), dim3(512), 0, s,
                           reinterpret_cast<const f128*>(dst + c_o), static_cast<uint4*>(e->rt4.p));
    static std::once_flag attr;  // engines step on threads of their own
    std::call_once(attr, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_batch_r0),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kR0Smem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_batch_fold),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kFoldSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_inner_r0_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kInnerGridSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_inner_fold_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kInnerGridSmem);
    });
    for (int u = 0; u < npos; ++u) {
        EnginePos& P = pos[u];
        // Binds (inner round 0 then reads D) and the final folds.
        for (size_t b : P.binds) {
            const auto bind_t0 = std::chrono::steady_clock::now();
            PiopJob* j = act[b];
            std::unique_ptr<BindScratch>& own = e->binds[j->m];
            if (!own) own.reset(new BindScratch);
            struct BindTime {
                std::chrono::steady_clock::time_point t0;
                double* ms;
                ~BindTime() {
                    *ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                }
            } bt{bind_t0, &bind_ms};
            if (j->devfs)
                TRY(bind_enqueue(j->m, j->fs_out(PiopJob::kOutRx), true, f128{0, 0}, j->dD.f(), s,
                                 &j->fs_state()->rho, own.get()));
            else
                TRY(bind_enqueue(j->m, reinterpret_cast<const f128*>(dst + rx_o[b]), true, j->rho, j->dD.f(), s,
                                 nullptr, own.get()));
        }
        for (size_t i : P.host_final) {
            PiopJob* j = act[i];
            f128* out = reinterpret_cast<f128*>(slot(i));
            if (j->phase == JobPhase::kOuterFinal) {
                bitz_outer* h = j->outer;
                hipLaunchKernelGGL(bitz::k_fold3, dim3(1), dim3(kBlock), 0, s, h->cur[0], h->cur[1], h->cur[2], out,
                                   out + 1, out + 2, (size_t)1, j->r);
            } else {
                hipLaunchKernelGGL(bitz::k_fold, dim3(1), dim3(kBlock), 0, s, j->icur[0], out, (size_t)1, j->r);
                hipLaunchKernelGGL(bitz::k_fold, dim3(1), dim3(kBlock), 0, s, j->icur[1], out + 1, (size_t)1, j->r);
            }
        }
        if (!P.o_r0.empty())
            hipLaunchKernelGGL(bitz::k_batch_r0, dim3((unsigned)P.o_r0_map.size()), dim3(bitz::kR0Block),
                               bitz::kR0Smem, s, reinterpret_cast<const bitz::RoundDesc*>(dst + P.or_o),
                               reinterpret_cast<const uint32_t*>(dst + P.orm_o));
        if (!P.o_fold.empty())
            hipLaunchKernelGGL(bitz::k_batch_fold, dim3((unsigned)P.o_fold_map.size()), dim3(bitz::kFoldBlock),
                               bitz::kFoldSmem, s, reinterpret_cast<const bitz::RoundDesc*>(dst + P.of_o),
                               reinterpret_cast<const uint32_t*>(dst + P.ofm_o));
        if (!P.i_r0.empty())
            hipLaunchKernelGGL(bitz::k_inner_r0_batch, dim3((unsigned)P.i_r0_map.size()), dim3(bitz::kInnerBlock),
                               bitz::kInnerGridSmem, s, reinterpret_cast<const bitz::InnerRoundDesc*>(dst + P.ir_o),
                               reinterpret_cast<const uint32_t*>(dst + P.irm_o));
        if (!P.i_bits.empty()) {
            // Every bit round of the position in one launch, each descriptor's
            // K picking its body.  Algorithmic bytes (bounds for
            // device-transcript proofs): round 0 reads D (32 B per pair) and
            // the bits; round K >= 1 reads 4 entries of D_{K-1} and writes 2
            // of D_K per output pair (and 2 of h_K when it writes them).
            double bits_bytes = 0;
            for (const auto& d : P.i_bits)
                bits_bytes += (d.k ? 96.0 + (d.wout ? 32.0 : 0.0) : 32.0) * (double)d.pairs +
                              (double)d.pairs * (2 << d.k) / 8;
            e->t_bits[u].begin(s, "inner_bits_batch", bits_bytes);
            hipLaunchKernelGGL(bitz::k_inner_bits_mixed, dim3((unsigned)P.i_bits_map.size()), dim3(bitz::kInnerBlock),
                               bitz::kInnerBitsSmem, s, reinterpret_cast<const bitz::InnerBitsDesc*>(dst + P.ib_o),
                               reinterpret_cast<const uint32_t*>(dst + P.ibm_o));
            e->t_bits[u].end(s);
        }
        if (!P.i_fold.empty()) {
            // Algorithmic bytes: per output pair, 4 + 4 inputs of D and h read
            // and 2 + 2 folded entries written.
            double fold_bytes = 0;
            for (const auto& d : P.i_fold) fold_bytes += 192.0 * (double)d.pairs;
            e->t_inner_fold[u].begin(s, "inner_fold_batch", fold_bytes);
            hipLaunchKernelGGL(bitz::k_inner_fold_batch, dim3((unsigned)P.i_fold_map.size()),
                               dim3(bitz::kInnerBlock), bitz::kInnerGridSmem, s,
                               reinterpret_cast<const bitz::InnerRoundDesc*>(dst + P.if_o),
                               reinterpret_cast<const uint32_t*>(dst + P.ifm_o));
            e->t_inner_fold[u].end(s);
        }
        if (!P.fs.empty())
            hipLaunchKernelGGL(bitz::k_piop_fs, dim3((unsigned)P.fs.size()), dim3(64), 0, s,
                               reinterpret_cast<const bitz::PiopFsDesc*>(dst + P.fs_o));
    }
    TRY(check_launch());
    if (slot_off[act.size()])
        HIP_TRY(hipMemcpyAsync(pin, e->slots.p, slot_off[act.size()] * 4, hipMemcpyDeviceToHost, s));
    for (size_t f = 0; f < finishing.size(); ++f) {
        PiopJob* j = act[finishing[f]];
        HIP_TRY(hipMemcpyAsync(pin + fin_back[f], j->fsb.p, j->fs_back_bytes(), hipMemcpyDeviceToHost, s));
    }
    e->t_step.end(s);
    HIP_TRY(hipEventRecord(e->step_done[b], s));
    for (size_t f = 0; f < finishing.size(); ++f) step.fin_back.emplace_back(act[finishing[f]], fin_back[f]);
    step.act = std::move(act);
    e->steps.push_back(std::move(step));
    e->next_buf ^= 1;
    prof_add_host("piop_engine_launch_kernels", kern_t0);
    if (g_prof.on) {
        std::lock_guard<std::mutex> lock(g_prof.mu);
        if (g_prof.only.empty()) {
            KernelStat& k = g_prof.stats["piop_engine_launch_binds"];
            k.ms += bind_ms;
            k.launches += 1;
        }
    }
    prof_add_host("piop_engine_host_launch", host_t0);
    return BITZ_OK;
}

extern "C" int32_t bitz_piop_engine_launch(bitz_piop_engine* e, size_t* launched) {
    if (!e) return fail(BITZ_ERR_ARG, "null pointer");
    if (e->running && std::this_thread::get_id() != e->worker.get_id())
        return fail(BITZ_ERR_STATE, "the engine steps in the background");
    if (!e->steps.empty()) return fail(BITZ_ERR_STATE, "a step is already in flight");
    return engine_launch_step(e, launched);
}

// Waits for the oldest step in flight, then the host's part: each finishing
// device-transcript proof's outputs, each host-transcript proof's round.
static int32_t engine_finish_step(bitz_piop_engine* e, size_t* in_flight) {
    std::vector<PiopJob*> act;
    EngineStep step;
    if (!e->steps.empty()) {
        step = std::move(e->steps.front());
        e->steps.pop_front();
        act = step.act;
    }
    const auto wait_t0 = std::chrono::steady_clock::now();
    if (!act.empty()) HIP_TRY(hipEventSynchronize(e->step_done[step.buf]));
    const auto host_t0 = std::chrono::steady_clock::now();
    if (!act.empty()) prof_add_host("piop_engine_wait", wait_t0);
    for (auto& t : e->t_inner_fold) t.collect();
    e->t_step.collect();
    for (auto& t : e->t_bits) t.collect();
    const char* pin = e->pin[step.buf].c();
    // Device-transcript proofs that ended in this step: their outputs, and
    // the duplex handed back to the caller's transcript.
    for (const auto& fb : step.fin_back) {
        PiopJob* j = fb.first;
        const char* b = pin + fb.second;
        bitz::PiopFsState st;
        memcpy(&st, b, sizeof st);
        j->t->t.keccak = st.ts;
        const bitz_f128* o = reinterpret_cast<const bitz_f128*>(b + sizeof st);
        auto take = [&](int w, bitz_f128* dst, size_t n) { memcpy(dst, o + j->out_off[w], n * sizeof(bitz_f128)); };
        take(PiopJob::kOutOuter, j->outer_rounds.data(), j->outer_rounds.size());
        take(PiopJob::kOutRx, j->rx.data(), j->rx.size());
        take(PiopJob::kOutClaims, j->outer_claims, 3);
        take(PiopJob::kOutInner, j->inner_rounds.data(), j->inner_rounds.size());
        take(PiopJob::kOutPoint, j->inner_point.data(), j->inner_point.size());
        take(PiopJob::kOutMatrixEval, &j->matrix_eval, 1);
        take(PiopJob::kOutFinal, &j->final_claim, 1);
        bitz_outer_free(j->outer);
        j->outer = nullptr;
        j->dlo.release();
        j->dhi.release();
        j->iP.release();
        j->iQ.release();
        j->dD.release();
        j->fsb.release();
        j->phase = JobPhase::kDone;
    }
    // Host-transcript proofs: the round polynomial into the transcript, the
    // next challenge.
    for (size_t i = 0; i < act.size(); ++i) {
        PiopJob* j = act[i];
        if (j->devfs) continue;
        const uint32_t* w = reinterpret_cast<const uint32_t*>(pin) + step.slot_off[i];
        Transcript& tr = j->t->t;
        switch (j->phase) {
            case JobPhase::kOuter: {
                bitz_outer* h = j->outer;
                const int k = j->round;
                if (k > 0) {
                    for (int u = 0; u < 3; ++u) h->cur[u] = j->pend[u];
                    h->where = (h->where == 1) ? 2 : 1;
                    h->len /= 2;
                    h->round += 1;
                }
                bitz_f128 raw[3], co[3];
                memcpy(raw, w, 48);
                outer_eq_coeffs(h, k, raw, j->q1, h->scale_big[k], co);
                j->r = recover_round(tr, j->claim, co, 3, j->outer_rounds.data() + 4 * k, j->rx.data() + k);
                j->round = k + 1;
                if (j->round == j->n) j->phase = JobPhase::kOuterFinal;
                break;
            }
            case JobPhase::kOuterFinal: {
                memcpy(j->outer_claims, w, 48);
                f128 cl[3] = {F(j->outer_claims[0]), F(j->outer_claims[1]), F(j->outer_claims[2])};
                tr.absorb_elems(0, cl, 3);  // sumcheck.rs:372
                // piop.rs:94-98: rho, the inner claim; D and the inner sumcheck follow.
                j->rho = tr.squeeze();
                j->claim = hostf::add(cl[0], hostf::add(hostf::mul(j->rho, cl[1]),
                                                        hostf::mul(hostf::mul(j->rho, j->rho), cl[2])));
                bitz_outer_free(j->outer);
                j->outer = nullptr;
                j->dlo.release();
                j->dhi.release();
                const size_t N = (size_t)1 << j->nc;
                j->ilen = N;
                j->ilive = N;
                j->icapP = std::max<size_t>(N / 2, 1);
                j->icapQ = std::max<size_t>(N / 4, 1);
                if ((j->status = j->dD.alloc(N * sizeof(f128))) || (j->status = j->iP.alloc(2 * j->icapP * sizeof(f128))) ||
                    (j->status = j->iQ.alloc(2 * j->icapQ * sizeof(f128)))) {
                    j->phase = JobPhase::kDone;
                    break;
                }
                j->icur[0] = j->dD.f();
                j->icur[1] = j->hv;
                j->round = 0;
                j->phase = JobPhase::kBind;
                break;
            }
            case JobPhase::kBind:
            case JobPhase::kInner: {
                const int k = j->round;
                const bool bit_round = j->hbits && k <= j->kb;
                if (k > 0) {
                    j->icur[0] = j->pend[0];
                    j->icur[1] = j->pend[1];
                    j->iwhere = (j->iwhere == 1) ? 2 : 1;
                    j->ilen /= 2;
                    j->ilive = (j->ilive + 1) / 2;
                }
                bitz_f128 co[2];
                unsigned long long live = 0;
                if (bit_round) {
                    const std::vector<f128> r = job_inner_point(j, k);
                    bits_round_coeffs(w, k, r.data(), co, &live);
                } else {
                    memcpy(co, w, 32);
                    memcpy(&live, w + 8, 8);
                }
                if (k == 0) {
                    // Round 0's live length (1 + the last index where D or h is
                    // nonzero): the fold rounds skip the zero tail.
                    j->ilive = std::max<size_t>(1, std::min<size_t>((size_t)live, j->ilen));
                    if (j->hbits) j->icur[1] = j->pend[1];  // h as elements if round 0 was the last bit round
                }
                j->r = recover_round(tr, j->claim, co, 2, j->inner_rounds.data() + 3 * k, j->inner_point.data() + k);
                j->round = k + 1;
                j->phase = j->round == j->nc ? JobPhase::kInnerFinal : JobPhase::kInner;
                break;
            }
            case JobPhase::kInnerFinal: {
                memcpy(&j->matrix_eval, w, 16);  // D(r_y); h(r_y) follows at w + 4
                j->final_claim = B(j->claim);
                j->iP.release();
                j->iQ.release();
                j->dD.release();
                j->phase = JobPhase::kDone;
                break;
            }
            case JobPhase::kDevDone:
            case JobPhase::kDone:
                break;
        }
    }
    if (in_flight) {
        size_t n = 0;
        for (auto& j : e->jobs) n += j->phase != JobPhase::kDone;
        *in_flight = n;
    }
    if (!act.empty()) prof_add_host("piop_engine_host_finish", host_t0);
    return BITZ_OK;
}

extern "C" int32_t bitz_piop_engine_finish(bitz_piop_engine* e, size_t* in_flight) {
    if (!e) return fail(BITZ_ERR_ARG, "null pointer");
    return engine_finish_step(e, in_flight);
}

extern "C" int32_t bitz_piop_engine_step(bitz_piop_engine* e, size_t* in_flight) {
    TRY(bitz_piop_engine_launch(e, nullptr));
    return bitz_piop_engine_finish(e, in_flight);
}

extern "C" int32_t bitz_piop_engine_done(bitz_piop_engine* e, uint64_t* tags, size_t cap, size_t* count) {
    if (!e || !count) return fail(BITZ_ERR_ARG, "null pointer");
    std::lock_guard<std::mutex> lock(e->mu);
    if (e->worker_rc) return fail(e->worker_rc, e->worker_err);
    size_t n = 0;
    // While the worker runs, only `finished` is the caller's to read.
    for (auto& j : e->running ? e->finished : e->jobs)
        if (j->phase == JobPhase::kDone) {
            if (tags && n < cap) tags[n] = j->tag;
            ++n;
        }
    *count = n;
    return BITZ_OK;
}

// The background thread: steps while proofs are in flight, sleeps otherwise.
// The background thread: steps while proofs are in flight, sleeps otherwise.
// When every active proof draws its challenges on the device (and kernel
// timing is off), the next step is queued before the one in flight is
// finished, so the device never waits for the host between steps.
static void piop_engine_worker(bitz_piop_engine* e) {
    auto launchable = [&] {
        for (auto& j : e->jobs)
            if (j->phase != JobPhase::kDone && j->phase != JobPhase::kDevDone) return true;
        return false;
    };
    auto all_devfs = [&] {
        for (auto& j : e->jobs)
            if (!j->devfs && j->phase != JobPhase::kDone) return false;
        return true;
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(e->mu);
            e->cv.wait(lock, [&] { return e->stop || !e->incoming.empty() || launchable() || !e->steps.empty(); });
            for (auto& j : e->incoming) e->jobs.push_back(std::move(j));
            e->incoming.clear();
            if (e->stop && !launchable() && e->steps.empty()) return;
        }
        // One step in flight: queueing a second behind it (the host building
        // step N + 1 while the device runs step N) measured 10-20 % fewer
        // blocks/s and a higher p50 under load; new proofs and finished ones
        // then wait a step longer to join and to free their slots.
        const size_t depth = 1;
        int32_t rc = BITZ_OK;
        if (e->steps.size() < depth && launchable()) rc = engine_launch_step(e, nullptr);
        if (!rc && !e->steps.empty() && (e->steps.size() >= depth || !launchable()))
            rc = engine_finish_step(e, nullptr);
        {
            std::lock_guard<std::mutex> lock(e->mu);
            if (rc) {
                e->worker_rc = rc;
                e->worker_err = bitz_last_error();
                (void)hipStreamSynchronize(e->so.s);
                e->steps.clear();
                for (auto& j : e->jobs) {
                    j->status = rc;
                    j->phase = JobPhase::kDone;
                }
            }
            for (auto& j : e->jobs)
                if (j && j->phase == JobPhase::kDone) e->finished.push_back(std::move(j));
            e->jobs.erase(std::remove(e->jobs.begin(), e->jobs.end(), nullptr), e->jobs.end());
        }
        e->cv.notify_all();
    }
}

extern "C" int32_t bitz_piop_engine_run(bitz_piop_engine* e) {
    if (!e) return fail(BITZ_ERR_ARG, "null pointer");
    if (e->running) return BITZ_OK;
    if (!e->steps.empty()) return fail(BITZ_ERR_STATE, "a step is in flight");
    e->running = true;
    e->worker = std::thread(piop_engine_worker, e);
    return BITZ_OK;
}

extern "C" int32_t bitz_piop_engine_wait(bitz_piop_engine* e, uint32_t timeout_us, size_t* finished) {
    if (!e || !finished) return fail(BITZ_ERR_ARG, "null pointer");
    std::unique_lock<std::mutex> lock(e->mu);
    e->cv.wait_for(lock, std::chrono::microseconds(timeout_us), [&] { return !e->finished.empty() || e->worker_rc; });
    *finished = e->finished.size();
    return BITZ_OK;
}

// Device bytes one proof holds in the engine: its own buffers and its outer
// sumcheck's, as allocated (size classes included); all are allocated when
// it is submitted and held until it is taken.
static uint64_t job_device_bytes(const PiopJob* j) {
    uint64_t b = j->dlo.bytes + j->dhi.bytes + j->dD.bytes + j->iP.bytes + j->iQ.bytes + j->fsb.bytes +
                 j->own_bits.bytes + j->own_prod.bytes;
    if (const bitz_outer* h = j->outer) {
        b += h->in.bytes + h->P.bytes + h->Q.bytes + h->acc.bytes + h->htabs4.bytes + h->rtabs4.bytes;
        for (int u = 0; u < 2; ++u) b += h->eqL[u].bytes + h->eqH[u].bytes;
    }
    return b;
}

extern "C" int32_t bitz_piop_engine_device_bytes(bitz_piop_engine* e, uint64_t tag, uint64_t* bytes) {
    if (!e || !bytes) return fail(BITZ_ERR_ARG, "null pointer");
    std::lock_guard<std::mutex> lock(e->mu);
    for (const auto* list : {&e->jobs, &e->incoming, &e->finished})
        for (const auto& j : *list)
            if (j->tag == tag) {
                *bytes = job_device_bytes(j.get());
                return BITZ_OK;
            }
    return fail(BITZ_ERR_ARG, "no proof with this tag");
}

extern "C" int32_t bitz_piop_engine_take(bitz_piop_engine* e, uint64_t tag, bitz_f128* outer_rounds,
                                         bitz_f128* outer_claims, bitz_f128* inner_rounds, bitz_f128* inner_point,
                                         bitz_f128* matrix_eval, bitz_f128* final_claim) {
    if (!e) return fail(BITZ_ERR_ARG, "null pointer");
    std::unique_lock<std::mutex> lock(e->mu);
    auto& list = e->running ? e->finished : e->jobs;
    for (size_t i = 0; i < list.size(); ++i) {
        PiopJob* j = list[i].get();
        if (j->tag != tag) continue;
        if (j->phase != JobPhase::kDone) return fail(BITZ_ERR_STATE, "proof still in flight");
        const int32_t st = j->status;
        if (st == BITZ_OK) {  // null outputs are not copied
            if (outer_rounds) memcpy(outer_rounds, j->outer_rounds.data(), j->outer_rounds.size() * sizeof(bitz_f128));
            if (outer_claims) memcpy(outer_claims, j->outer_claims, 3 * sizeof(bitz_f128));
            if (inner_rounds) memcpy(inner_rounds, j->inner_rounds.data(), j->inner_rounds.size() * sizeof(bitz_f128));
            if (inner_point) memcpy(inner_point, j->inner_point.data(), j->inner_point.size() * sizeof(bitz_f128));
            if (matrix_eval) *matrix_eval = j->matrix_eval;
            if (final_claim) *final_claim = j->final_claim;
        }
        list.erase(list.begin() + (std::ptrdiff_t)i);
        return st == BITZ_OK ? BITZ_OK : fail(st, "proof failed");
    }
    for (const auto* other : {&e->jobs, &e->incoming})
        for (const auto& j : *other)
            if (j && j->tag == tag) return fail(BITZ_ERR_STATE, "proof still in flight");  // tags only: set at submit
    return fail(BITZ_ERR_ARG, "no such tag");
}


// ================================================================ M^T =====
struct bitz_mt {
    size_t row_count = 0, cols = 0, nnz = 0;
    Dev offsets, rows;
    Dev long_cols;          // columns of more than kMtLongColumn entries
    uint32_t n_long = 0;
    StreamOwner so;  // last: destroyed (and synchronised) before the buffers
};

extern "C" int32_t bitz_mt_prepare(bitz_mt** out, size_t row_count, size_t columns,
                                   const uint32_t* column_offsets, const uint32_t* row_indices) {
    if (!out || !column_offsets) return fail(BITZ_ERR_ARG, "null pointer");
    const size_t nnz = column_offsets[columns];
    if (nnz && !row_indices) return fail(BITZ_ERR_ARG, "null row indices");
    for (size_t k = 0; k < nnz; ++k)
        if (row_indices[k] >= row_count) return fail(BITZ_ERR_ARG, "row index out of range");
    for (size_t c = 0; c < columns; ++c)
        if (column_offsets[c] > column_offsets[c + 1]) return fail(BITZ_ERR_ARG, "offsets not monotone");
    auto* h = new bitz_mt;
    auto cleanup = [&](int32_t c) {
        delete h;
        return c;
    };
    int32_t rc;
    if ((rc = h->so.create())) return cleanup(rc);
    h->row_count = row_count;
    h->cols = columns;
    h->nnz = nnz;
    if ((rc = h->offsets.alloc((columns + 1) * 4))) return cleanup(rc);
    if ((rc = h->rows.alloc(std::max<size_t>(nnz, 1) * 4))) return cleanup(rc);
    if (hipMemcpy(h->offsets.p, column_offsets, (columns + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        (nnz && hipMemcpy(h->rows.p, row_indices, nnz * 4, hipMemcpyHostToDevice) != hipSuccess))
        return cleanup(fail(BITZ_ERR_DEVICE, "upload failed"));
    std::vector<uint32_t> longs;
    for (size_t c = 0; c < columns; ++c)
        if (column_offsets[c + 1] - column_offsets[c] > bitz::kMtLongColumn) longs.push_back((uint32_t)c);
    h->n_long = (uint32_t)longs.size();
    if ((rc = h->long_cols.alloc(std::max<size_t>(longs.size(), 1) * 4))) return cleanup(rc);
    if (!longs.empty() && hipMemcpy(h->long_cols.p, longs.data(), longs.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(BITZ_ERR_DEVICE, "upload failed"));
    *out = h;
    return BITZ_OK;
}

extern "C" int32_t bitz_mt_apply(bitz_mt* h, const bitz_f128* challenges, size_t n, bitz_f128* out,
                                 uint32_t flags) {
    if (!h || !out || (n && !challenges)) return fail(BITZ_ERR_ARG, "null pointer");
    // matrix_transpose.rs:346-351: one challenge per row of M.
    if (n != h->row_count) return fail(BITZ_ERR_ARG, "challenge count != row count");
    hipStream_t s = h->so.s;
    Dev dc, dout;
    if (n)
        TRY(upload(dc, challenges, n, flags, s));
    else
        TRY(dc.alloc(sizeof(f128)));
    f128* target = reinterpret_cast<f128*>(out);
    if (!(flags & BITZ_DEVICE)) {
        TRY(dout.alloc(h->cols * sizeof(f128)));
        target = dout.f();
    }
    // Timed as "mt_apply": the CSC (4 B per column and nonzero), every
    // challenge read once and the output written once.
    Timed tm;
    tm.begin(s, "mt_apply", 4.0 * (double)(h->cols + 1 + h->nnz) + 16.0 * (double)(n + h->cols));
    hipLaunchKernelGGL(bitz::k_mt_apply, dim3(grid_for(h->cols)), dim3(kBlock), 0, s, h->offsets.u32(),
                       h->rows.u32(), dc.f(), target, h->cols);
    if (h->n_long)
        hipLaunchKernelGGL(bitz::k_mt_apply_long, dim3(h->n_long), dim3(kBlock), 0, s, h->offsets.u32(),
                           h->rows.u32(), dc.f(), h->long_cols.u32(), target);
    tm.end(s);
    TRY(check_launch());
    if (flags & BITZ_DEVICE)
        HIP_TRY(hipStreamSynchronize(s));
    else
        TRY(download(out, target, h->cols, 0, s));
    tm.collect();
    return BITZ_OK;
}

extern "C" int32_t bitz_mt_transpose_inner(bitz_mt* h, const bitz_f128* row_w, size_t n_row,
                                           const bitz_f128* col_w, size_t n_col, bitz_f128* out,
                                           uint32_t flags) {
    if (!h || !out || !row_w || !col_w) return fail(BITZ_ERR_ARG, "null pointer");
    // matrix_transpose.rs:386-392: power-of-two rows per column covering M.
    if (n_row == 0 || (n_row & (n_row - 1)) || (n_col && n_row * n_col < h->row_count) || n_col == 0)
        return fail(BITZ_ERR_ARG, "weight count mismatch");
    const int log_rows = log2_exact(n_row);
    // Every row index must land inside col_w.
    if (((h->row_count ? h->row_count - 1 : 0) >> log_rows) >= n_col)
        return fail(BITZ_ERR_ARG, "weight count mismatch");
    hipStream_t s = h->so.s;
    Dev dr, dc, dout;
    TRY(upload(dr, row_w, n_row, flags, s));
    TRY(upload(dc, col_w, n_col, flags, s));
    f128* target = reinterpret_cast<f128*>(out);
    if (!(flags & BITZ_DEVICE)) {
        TRY(dout.alloc(h->cols * sizeof(f128)));
        target = dout.f();
    }
    // Timed as "mt_transpose_inner": the CSC, both weight tables read once
    // and the output written once.
    Timed tm;
    tm.begin(s, "mt_transpose_inner",
             4.0 * (double)(h->cols + 1 + h->nnz) + 16.0 * (double)(n_row + n_col + h->cols));
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_mt_transpose_inner),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_mt_inner_long),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
        attr = true;
    }
    hipLaunchKernelGGL(bitz::k_mt_transpose_inner, dim3(grid_for(h->cols)), dim3(kBlock), bitz::kLaneSmem256, s,
                       h->offsets.u32(), h->rows.u32(), dr.f(), log_rows, dc.f(), target, h->cols);
    if (h->n_long)
        hipLaunchKernelGGL(bitz::k_mt_inner_long, dim3(h->n_long), dim3(kBlock), bitz::kLaneSmem256, s,
                           h->offsets.u32(), h->rows.u32(), dr.f(), log_rows, dc.f(), h->long_cols.u32(), target);
    tm.end(s);
    TRY(check_launch());
    if (flags & BITZ_DEVICE)
        HIP_TRY(hipStreamSynchronize(s));
    else
        TRY(download(out, target, h->cols, 0, s));
    tm.collect();
    return BITZ_OK;
}

extern "C" void bitz_mt_free(bitz_mt* h) { delete h; }

// ================================================================== GKR ===
struct bitz_gkr {
    int log_leaves = 0, log_groups = 0;
    Dev store, eq, eq_scratch, point, acc;
    std::vector<size_t> off;  // offset of layer j (0 = leaves)
    Pinned host;
    bool consumed = false;
    Timed tm;
    StreamOwner so;  // last: destroyed (and synchronised) before the buffers
};

constexpr size_t kSmallTreeSmem = sizeof(uint4) * ((size_t)1 << bitz::kGkrSmallTreeLog) + bitz::kLaneSmem256;

// The product tree over layer 0.  `fill` writes the leaves into the store;
// with `src` a device table instead, the first layer reads it and copies it
// into the store itself (the leaves are read once).
template <class Fill>
static int32_t gkr_build(bitz_gkr** out, uint32_t log_leaves, uint32_t log_groups, const f128* src, Fill fill,
                         bitz_f128* output, uint32_t flags) {
    auto* h = new bitz_gkr;
    auto cleanup = [&](int32_t c) {
        delete h;
        return c;
    };
    int32_t rc;
    if ((rc = h->so.create())) return cleanup(rc);
    hipStream_t s = h->so.s;
    h->log_leaves = (int)log_leaves;
    h->log_groups = (int)log_groups;
    size_t total = 0;
    for (int j = 0; j <= (int)(log_leaves - log_groups); ++j) {
        h->off.push_back(total);
        total += (size_t)1 << (log_leaves - j);
    }
    if ((rc = h->store.alloc(total * sizeof(f128)))) return cleanup(rc);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_layer_lane),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_layers_small),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSmallTreeSmem);
        attr = true;
    }
    // Layers are contiguous: layer j+1 starts where layer j ends. Large layers
    // take one grid-wide launch each; from 2^kGkrSmallTreeLog elements down
    // one workgroup builds the rest.
    const int depth = (int)(log_leaves - log_groups);
    const int j_small = std::max(0, (int)log_leaves - bitz::kGkrSmallTreeLog);
    const bool fused_copy = src && depth > 0 && j_small > 0;
    if (!fused_copy && (rc = fill(h, s))) return cleanup(rc);
    // Timed as "gkr_layers": every layer read once, the next written once.
    Timed tl;
    tl.begin(s, "gkr_layers", 24.0 * (double)(((size_t)1 << log_leaves) - ((size_t)1 << log_groups)));
    for (int j = 0; j < std::min(depth, j_small); ++j) {
        const size_t half = (size_t)1 << (log_leaves - j - 1);
        const bool copy = fused_copy && j == 0;
        hipLaunchKernelGGL(bitz::k_gkr_layer_lane, dim3(grid_for(half)), dim3(bitz::kStageBlock),
                           bitz::kLaneSmem256, s, copy ? src : h->store.f() + h->off[j], copy ? h->store.f() : nullptr,
                           h->store.f() + h->off[j + 1], half);
    }
    if (depth > j_small)
        hipLaunchKernelGGL(bitz::k_gkr_layers_small, dim3(1), dim3(bitz::kStageBlock), kSmallTreeSmem, s,
                           h->store.f() + h->off[j_small], (int)log_leaves - j_small, depth - j_small);
    tl.end(s);
    if ((rc = check_launch())) return cleanup(rc);
    if ((rc = download(output, h->store.f() + h->off.back(), (size_t)1 << log_groups, flags, s)))
        return cleanup(rc);
    tl.collect();
    if ((rc = h->host.alloc(64))) return cleanup(rc);
    *out = h;
    return BITZ_OK;
}

extern "C" int32_t bitz_gkr_build(bitz_gkr** out, const bitz_f128* leaves, uint32_t log_leaves,
                                  uint32_t log_groups, bitz_f128* output, uint32_t flags) {
    if (!out || !leaves || !output) return fail(BITZ_ERR_ARG, "null pointer");
    if (log_groups > log_leaves || log_leaves > 32) return fail(BITZ_ERR_ARG, "bad tree shape");
    const bool dev = (flags & BITZ_DEVICE) != 0;
    auto fill = [&](bitz_gkr* h, hipStream_t s) -> int32_t {
        if (hipMemcpyAsync(h->store.p, leaves, ((size_t)1 << log_leaves) * sizeof(f128),
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s) != hipSuccess)
            return fail(BITZ_ERR_DEVICE, "leaf upload failed");
        return BITZ_OK;
    };
    return gkr_build(out, log_leaves, log_groups, dev ? reinterpret_cast<const f128*>(leaves) : nullptr, fill,
                     output, flags);
}

extern "C" int32_t bitz_gkr_build_bittable(bitz_gkr** out, const bitz_f128* images, uint32_t log_rows,
                                           const uint64_t* packed, uint32_t log_columns, bitz_f128* output,
                                           uint32_t flags) {
    if (!out || !images || !packed || !output) return fail(BITZ_ERR_ARG, "null pointer");
    // LutLayers::new (crates/gkr/src/lib.rs:274-302) needs a word of nibbles per lookup table.
    if (log_columns < 4) return fail(BITZ_ERR_ARG, "the table needs at least 16 trees (s >= 4)");
    if (log_rows + log_columns > 32) return fail(BITZ_ERR_ARG, "bad table shape");
    const bool dev = (flags & BITZ_DEVICE) != 0;
    const size_t rows = (size_t)1 << log_rows, words = std::max<size_t>(((size_t)1 << (log_rows + log_columns)) / 64, 1);
    Dev dimg, dbits;
    // Timed as "bittable_leaves": the bits and images read, 16 B per leaf written.
    Timed te;
    auto fill = [&](bitz_gkr* h, hipStream_t s) -> int32_t {
        const f128* img = reinterpret_cast<const f128*>(images);
        const uint64_t* bits = packed;
        if (!dev) {
            TRY(upload(dimg, images, rows, 0, s));
            TRY(dbits.alloc(words * 8));
            HIP_TRY(hipMemcpyAsync(dbits.p, packed, words * 8, hipMemcpyHostToDevice, s));
            img = dimg.f();
            bits = static_cast<const uint64_t*>(dbits.p);
        }
        te.begin(s, "bittable_leaves", 16.0 * (double)(rows + (rows << log_columns)) + 8.0 * (double)words);
        hipLaunchKernelGGL(bitz::k_bittable_leaves, dim3(grid_for(rows << log_columns)), dim3(kBlock), 0, s, bits,
                           img, (int)log_rows, (int)log_columns, h->store.f());
        te.end(s);
        return check_launch();
    };
    const int32_t rc = gkr_build(out, log_rows + log_columns, log_columns, nullptr, fill, output, flags);
    // the staged table and images must outlive the expansion kernel
    if (rc == BITZ_OK && !dev) HIP_TRY(hipStreamSynchronize((*out)->so.s));
    if (rc == BITZ_OK) te.collect();  // gkr_build synchronised the stream
    return rc;
}

extern "C" int32_t bitz_gkr_layer(bitz_gkr* h, uint32_t j, bitz_f128* out, size_t cap, size_t* len) {
    if (!h || !len) return fail(BITZ_ERR_ARG, "null pointer");
    if (j >= h->off.size()) return fail(BITZ_ERR_ARG, "no such layer");
    const size_t n = (size_t)1 << (h->log_leaves - (int)j);
    *len = n;
    if (out) {
        HIP_TRY(hipMemcpyAsync(out, h->store.f() + h->off[j], std::min(cap, n) * sizeof(f128),
                               hipMemcpyDeviceToHost, h->so.s));
        HIP_TRY(hipStreamSynchronize(h->so.s));
    }
    return BITZ_OK;
}

// Whether provers with a Keccak transcript draw their challenges on the
// device (bitz_set_device_fiat_shamir); on by default (g_device_fs, above).

extern "C" int32_t bitz_set_engine_positions(int32_t positions) {
    if (positions < 0 || positions > kEngineMaxPositions) return fail(BITZ_ERR_ARG, "positions out of range");
    g_engine_positions = positions;
    return BITZ_OK;
}

extern "C" int32_t bitz_set_device_fiat_shamir(int32_t on) {
    g_device_fs = on != 0;
    return BITZ_OK;
}

// gpgkr_prove with the transcript on the device (a Keccak transcript, no
// grinding): the layers in the reference's order, each layer's eq table from
// its point on the device, the grid-wide rounds each with its challenge
// (k_gkr_round_fs, whose last workgroup draws it, or k_gkr_round_batch then
// k_gkr_fs_round for the large ones), the last rounds and the layer's end in
// k_gkr_tail_fs.
// Nothing waits for the host until the end, when the transcript state, the
// leaf point, the claim and the prover messages (appended to the narg string
// in their order) come back.
static int32_t gkr_prove_device_fs(bitz_gkr* h, bitz_transcript* t, const std::vector<f128>& top,
                                   bitz_f128* leaf_point, bitz_f128* leaf_claim) {
    hipStream_t s = h->so.s;
    constexpr int T = bitz::kGkrTailMax;
    const int layers = h->log_leaves - h->log_groups;
    // Layer l (0 = the top) has n = log_groups + l coordinates; its point is
    // at pt_off[l], the leaf point at pt_off[layers].
    std::vector<size_t> pt_off(layers + 1), msg_off(layers), slot_off(layers);
    size_t n_pts = 0, n_msgs = 0, n_slots = 0, eq_scratch = 0;
    for (int l = 0; l <= layers; ++l) {
        pt_off[l] = n_pts;
        n_pts += (size_t)(h->log_groups + l);
    }
    for (int l = 0; l < layers; ++l) {
        const int n = h->log_groups + l, k = n - 1;
        msg_off[l] = n_msgs;
        n_msgs += 2 * ((size_t)n + 1);
        slot_off[l] = n_slots;
        n_slots += (size_t)n;
        if (k > bitz::kGkrEqSmallMax) eq_scratch = std::max(eq_scratch, ((size_t)1 << (k / 2)) + ((size_t)1 << (k - k / 2)));
    }
    Dev P, M, st, rtab, eqs, desc, done;
    TRY(P.alloc(n_pts * sizeof(f128)));
    TRY(done.alloc(std::max<size_t>(n_slots, 1) * sizeof(uint32_t)));
    TRY(M.alloc(n_msgs * sizeof(f128)));
    TRY(st.alloc(sizeof(bitz::GkrFsState)));
    TRY(rtab.alloc(kConstTable4Entries * sizeof(uint4)));
    TRY(eqs.alloc(std::max<size_t>(eq_scratch, 1) * sizeof(f128)));
    TRY(h->acc.alloc(std::max<size_t>(n_slots, 1) * 8 * sizeof(uint32_t)));
    TRY(h->eq.alloc(((size_t)1 << std::max(h->log_leaves - 2, 0)) * sizeof(f128)));
    // Descriptors of every launch: pointers only, so all are known up front.
    struct LayerPlan {
        size_t eq_items = 0, eq_items_n = 0, combine = 0;
        bool split = false;
        uint32_t combine_blocks = 0;
        size_t rounds = 0;  // first grid-wide round descriptor
        int grid_rounds = 0;
        size_t tail = 0;
    };
    std::vector<LayerPlan> plan(layers);
    std::vector<bitz::GkrEqItem> eq_items;
    std::vector<bitz::GkrCombine> combines;
    std::vector<bitz::GkrRound> rounds;
    std::vector<bitz::GkrTail> tails;
    uint32_t max_blocks = 1;
    for (int l = 0; l < layers; ++l) {
        const int n = h->log_groups + l, k = n - 1;
        const int j = layers - 1 - l;  // storage layer (0 = leaves): the children of layer l's nodes
        f128* left = h->store.f() + h->off[j];
        f128* right = left + ((size_t)1 << n);
        const f128* pt = P.f() + pt_off[l];
        LayerPlan& lp = plan[l];
        lp.eq_items = eq_items.size();
        if (k <= bitz::kGkrEqSmallMax) {
            eq_items.push_back({pt, k, h->eq.f()});
        } else {
            const int kl = k / 2, kh = k - kl;
            f128* low = eqs.f();
            f128* high = low + ((size_t)1 << kl);
            eq_items.push_back({pt, kl, low});
            eq_items.push_back({pt + kl, kh, high});
            bitz::GkrCombine c{};
            c.low = low;
            c.high = high;
            c.out = h->eq.f();
            c.n = (size_t)1 << k;
            c.low_bits = kl;
            c.blocks = grid_for(c.n);
            c.block_base = 0;
            lp.split = true;
            lp.combine = combines.size();
            lp.combine_blocks = c.blocks;
            max_blocks = std::max(max_blocks, c.blocks);
            combines.push_back(c);
        }
        lp.eq_items_n = eq_items.size() - lp.eq_items;
        lp.rounds = rounds.size();
        for (int i = n - 1; i > T; --i) {
            bitz::GkrRound g{};
            g.left = left;
            g.right = right;
            g.eq = h->eq.f();
            g.h = (size_t)1 << i;
            g.first = i == n - 1;
            if (!g.first) g.rt4 = static_cast<const uint4*>(rtab.p);
            g.acc = h->acc.u32() + 8 * (slot_off[l] + (size_t)(n - 1 - i));
            g.z = pt + i;
            g.blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>((g.h + bitz::kStageBlock - 1) / bitz::kStageBlock, 2048));
            g.block_base = 0;
            max_blocks = std::max(max_blocks, g.blocks);
            rounds.push_back(g);
            ++lp.grid_rounds;
        }
        bitz::GkrTail tl{};
        tl.left = left;
        tl.right = right;
        tl.eq = h->eq.f();
        tl.n = n;
        tl.i0 = std::min(n - 1, T);
        tl.rtabs4 = tl.i0 < n - 1 ? static_cast<const uint4*>(rtab.p) : nullptr;
        lp.tail = tails.size();
        tails.push_back(tl);
    }
    size_t off = 0;
    auto place = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        const size_t o = off;
        off += bytes;
        return o;
    };
    const size_t eqi_o = place(eq_items.size() * sizeof(bitz::GkrEqItem));
    const size_t cmb_o = place(combines.size() * sizeof(bitz::GkrCombine));
    const size_t rd_o = place(rounds.size() * sizeof(bitz::GkrRound));
    const size_t tl_o = place(tails.size() * sizeof(bitz::GkrTail));
    const size_t map_o = place((size_t)max_blocks * 4);  // all zeros: every launch has one descriptor
    std::vector<char> stage(off, 0);
    memcpy(stage.data() + eqi_o, eq_items.data(), eq_items.size() * sizeof(bitz::GkrEqItem));
    memcpy(stage.data() + cmb_o, combines.data(), combines.size() * sizeof(bitz::GkrCombine));
    memcpy(stage.data() + rd_o, rounds.data(), rounds.size() * sizeof(bitz::GkrRound));
    memcpy(stage.data() + tl_o, tails.data(), tails.size() * sizeof(bitz::GkrTail));
    TRY(desc.alloc(std::max<size_t>(off, 16)));
    bitz::GkrFsState init{};
    init.ts = t->t.keccak;
    init.eq_bound = hostf::one();
    HIP_TRY(hipMemcpyAsync(desc.p, stage.data(), off, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(P.p, top.data(), top.size() * sizeof(f128), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(st.p, &init, sizeof init, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(h->acc.p, 0, h->acc.bytes, s));
    HIP_TRY(hipMemsetAsync(done.p, 0, done.bytes, s));
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_round_fs),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kGkrRoundSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_tail_fs),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kGkrTailSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_round_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kGkrRoundSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_eq_items),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_eq_combine_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
        attr = true;
    }
    const char* dp = static_cast<const char*>(desc.p);
    const uint32_t* zmap = reinterpret_cast<const uint32_t*>(dp + map_o);
    auto* state = static_cast<bitz::GkrFsState*>(st.p);
    h->tm.begin(s, "gkr_prove_device_fs", 0);
    for (int l = 0; l < layers; ++l) {
        const int n = h->log_groups + l;
        const LayerPlan& lp = plan[l];
        hipLaunchKernelGGL(bitz::k_gkr_eq_items, dim3((unsigned)lp.eq_items_n), dim3(bitz::kStageBlock),
                           bitz::kLaneSmem256, s,
                           reinterpret_cast<const bitz::GkrEqItem*>(dp + eqi_o) + lp.eq_items);
        if (lp.split)
            hipLaunchKernelGGL(bitz::k_gkr_eq_combine_batch, dim3(lp.combine_blocks), dim3(bitz::kStageBlock),
                               bitz::kLaneSmem256, s,
                               reinterpret_cast<const bitz::GkrCombine*>(dp + cmb_o) + lp.combine, zmap);
        for (int g = 0; g < lp.grid_rounds; ++g) {
            const bitz::GkrRound& R = rounds[lp.rounds + g];
            const int i = n - 1 - g;
            bitz::GkrFsRound fr{};
            fr.st = state;
            fr.z = P.f() + pt_off[l] + i;
            fr.r_out = P.f() + pt_off[l + 1] + i;
            fr.rtab_out = static_cast<uint4*>(rtab.p);
            fr.msg_out = M.f() + msg_off[l] + 2 * (size_t)(n - 1 - i);
            const bitz::GkrRound* rd = reinterpret_cast<const bitz::GkrRound*>(dp + rd_o) + lp.rounds + g;
            if (R.blocks <= bitz::kGkrFusedMaxBlocks) {
                hipLaunchKernelGGL(bitz::k_gkr_round_fs, dim3(R.blocks), dim3(bitz::kStageBlock), bitz::kGkrRoundSmem,
                                   s, rd, fr, done.u32() + slot_off[l] + (size_t)(n - 1 - i));
            } else {
                hipLaunchKernelGGL(bitz::k_gkr_round_batch, dim3(R.blocks), dim3(bitz::kStageBlock),
                                   bitz::kGkrRoundSmem, s, rd, zmap);
                hipLaunchKernelGGL(bitz::k_gkr_fs_round, dim3(1), dim3(bitz::kStageBlock), 0, s, rd, fr);
            }
        }
        bitz::GkrFsTail ft{};
        ft.st = state;
        ft.point = P.f() + pt_off[l];
        ft.next = P.f() + pt_off[l + 1];
        ft.msgs = M.f() + msg_off[l];
        hipLaunchKernelGGL(bitz::k_gkr_tail_fs, dim3(1), dim3(bitz::kStageBlock), bitz::kGkrTailSmem, s,
                           reinterpret_cast<const bitz::GkrTail*>(dp + tl_o) + lp.tail, ft);
        TRY(check_launch());
    }
    h->tm.end(s);
    bitz::GkrFsState fin{};
    std::vector<f128> msgs(n_msgs), leaf((size_t)h->log_leaves);
    HIP_TRY(hipMemcpyAsync(&fin, st.p, sizeof fin, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(msgs.data(), M.p, n_msgs * sizeof(f128), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(leaf.data(), P.f() + pt_off[layers], leaf.size() * sizeof(f128), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    h->tm.collect();
    // The host transcript continues from the device's state; the prover
    // messages join its narg string in the order they were absorbed.
    t->t.keccak = fin.ts;
    const uint8_t* mb = reinterpret_cast<const uint8_t*>(msgs.data());
    t->t.narg.insert(t->t.narg.end(), mb, mb + n_msgs * sizeof(f128));
    for (size_t i = 0; i < leaf.size(); ++i) leaf_point[i] = B(leaf[i]);
    *leaf_claim = B(fin.claim);
    return BITZ_OK;
}

extern "C" int32_t bitz_gkr_prove(bitz_gkr* h, bitz_transcript* t, const bitz_f128* point,
                                  uint32_t point_len, uint32_t security_bits, bitz_f128* leaf_point,
                                  bitz_f128* leaf_claim) {
    if (!h || !t || !leaf_point || !leaf_claim || (point_len && !point))
        return fail(BITZ_ERR_ARG, "null pointer");
    if (security_bits != 100 && security_bits != 128)
        return fail(BITZ_ERR_ARG, "security_bits is a SecurityLevel: 100 or 128");
    // lib.rs:144: grinding_bits(3) before every cubic round's challenge.
    static const uint8_t kCubicLabel[] = "gkr/cubic/v1";  // lib.rs:10
    const uint32_t grind_bits = bitz::grinding_bits(security_bits, 3);
    // lib.rs:39-43: one coordinate per tree bit; EqTables needs at least one.
    if ((int)point_len != h->log_groups || point_len == 0)
        return fail(BITZ_ERR_POINT_LEN, "the output has one value per tree, so the point has a coordinate per tree bit");
    if (h->consumed) return fail(BITZ_ERR_STATE, "layers already consumed");
    h->consumed = true;
    if (t->t.type == Transcript::Type::kKeccak && grind_bits == 0 && g_device_fs) {
        std::vector<f128> top(point_len);
        for (uint32_t i = 0; i < point_len; ++i) top[i] = F(point[i]);
        return gkr_prove_device_fs(h, t, top, leaf_point, leaf_claim);
    }
    hipStream_t s = h->so.s;
    const int layers = h->log_leaves - h->log_groups;
    int total_rounds = 0;
    for (int n = h->log_groups; n < h->log_leaves; ++n) total_rounds += n;
    TRY(h->acc.alloc((size_t)(total_rounds + 1) * 8 * sizeof(uint32_t)));
    HIP_TRY(hipMemsetAsync(h->acc.p, 0, h->acc.bytes, s));
    TRY(h->eq.alloc(((size_t)1 << std::max(h->log_leaves - 2, 0)) * sizeof(f128)));
    TRY(h->point.alloc((size_t)(h->log_leaves + 1) * sizeof(f128)));
    std::vector<f128> pt(point_len);
    for (uint32_t i = 0; i < point_len; ++i) pt[i] = F(point[i]);
    int slot = 0;
    f128 claim{0, 0};
    for (int j = layers - 1; j >= 0; --j) {
        const int n = (int)pt.size();
        f128* left = h->store.f() + h->off[j];
        f128* right = left + ((size_t)1 << n);
        // EqTables (lib.rs:663-693): only the largest, eq(z_<n-1, .), is built;
        // each round halves it for the next.
        if (n > 1) HIP_TRY(hipMemcpy(h->point.p, pt.data(), (n - 1) * sizeof(f128), hipMemcpyHostToDevice));
        TRY(eq_build(h->point.f(), n - 1, h->eq.f(), h->eq_scratch, s));
        std::vector<f128> next(n + 1);
        f128 eq_bound = hostf::one();
        f128 r_prev{0, 0};
        for (int i = n - 1; i >= 0; --i) {
            const f128 z = pt[i];
            const int at_one = hostf::is_zero(z);  // sends_one (lib.rs:622-624)
            uint32_t* acc = h->acc.u32() + 8 * slot;
            if (i == n - 1) {
                const size_t half = (size_t)1 << i;
                h->tm.begin(s, "gkr_sums", 80.0 * (double)half);
                hipLaunchKernelGGL(bitz::k_gkr_sums, dim3(grid_for(half)), dim3(kBlock), 0, s, left, right,
                                   h->eq.f(), half, at_one, acc);
            } else {
                const size_t hh = (size_t)1 << i;
                h->tm.begin(s, "gkr_bind_next", 240.0 * (double)hh);
                hipLaunchKernelGGL(bitz::k_gkr_bind_next, dim3(grid_for(hh)), dim3(kBlock), 0, s, left, right,
                                   h->eq.f(), hh, r_prev, at_one, acc);
            }
            h->tm.end(s);
            TRY(check_launch());
            HIP_TRY(hipMemcpyAsync(h->host.p, acc, 32, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            h->tm.collect();
            ++slot;
            f128 sums[2];
            memcpy(sums, h->host.p, 32);
            // lib.rs:142: prover_message([eq_bound * sum_b, eq_bound * sum_inf]).
            f128 msg[2] = {hostf::mul(eq_bound, sums[0]), hostf::mul(eq_bound, sums[1])};
            t->t.absorb_elems(1, msg, 2);
            t->t.grind(kCubicLabel, sizeof kCubicLabel - 1, grind_bits);
            const f128 r = t->t.squeeze();
            next[i] = r;
            eq_bound = hostf::mul(eq_bound, f128{1 ^ r.lo ^ z.lo, r.hi ^ z.hi});  // eq_factor
            r_prev = r;
        }
        hipLaunchKernelGGL(bitz::k_gkr_bind, dim3(1), dim3(kBlock), 0, s, left, right, (size_t)1, r_prev);
        TRY(check_launch());
        f128 lr[2];
        HIP_TRY(hipMemcpyAsync(h->host.p, left, 16, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(static_cast<char*>(h->host.p) + 16, right, 16, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        memcpy(lr, h->host.p, 32);
        t->t.absorb_elems(1, lr, 2);  // lib.rs:154
        const f128 tt = t->t.squeeze();
        next[n] = tt;
        claim = hostf::add(lr[0], hostf::mul(tt, hostf::add(lr[1], lr[0])));
        pt = next;
    }
    for (size_t i = 0; i < pt.size(); ++i) leaf_point[i] = B(pt[i]);
    *leaf_claim = B(claim);
    return BITZ_OK;
}

// The same proof with the challenges known in advance (a transcript whose
// challenges do not depend on the messages, such as the seeded stream):
// every layer's point follows from them, so the layers are independent and
// run side by side: their eq tables in two launches, the grid-wide rounds of
// the large layers one launch per round index, every layer's last rounds in
// one launch of LDS tails, then one read-back.  challenges: per layer from
// the top, the n round challenges in squeeze order (coordinate n-1 first),
// then t.  messages receives [eq_bound sum_b, eq_bound sum_inf] per round in
// the same order, layer_evals [L(r), R(r)] per layer.
extern "C" int32_t bitz_gkr_prove_with_challenges(bitz_gkr* h, const bitz_f128* point, uint32_t point_len,
                                                  const bitz_f128* challenges, size_t challenge_count,
                                                  bitz_f128* messages, bitz_f128* layer_evals,
                                                  bitz_f128* leaf_point, bitz_f128* leaf_claim) {
    if (!h || !leaf_point || !leaf_claim || (point_len && !point)) return fail(BITZ_ERR_ARG, "null pointer");
    if ((int)point_len != h->log_groups || point_len == 0)
        return fail(BITZ_ERR_POINT_LEN, "the output has one value per tree, so the point has a coordinate per tree bit");
    const int layers = h->log_leaves - h->log_groups;
    size_t need = 0, total_rounds = 0;
    for (int n = h->log_groups; n < h->log_leaves; ++n) {
        need += (size_t)n + 1;
        total_rounds += (size_t)n;
    }
    if (challenge_count != need) return fail(BITZ_ERR_ARG, "challenge count must be sum over layers of (n + 1)");
    if (need && (!challenges || !layer_evals || (total_rounds && !messages))) return fail(BITZ_ERR_ARG, "null pointer");
    if (h->consumed) return fail(BITZ_ERR_STATE, "layers already consumed");
    h->consumed = true;
    hipStream_t s = h->so.s;
    constexpr int T = bitz::kGkrTailMax;
    // Every layer's point: the top one given, then (r_0 .. r_{n-1}, t).
    std::vector<std::vector<f128>> pts(layers + 1);
    pts[0].resize(point_len);
    for (uint32_t i = 0; i < point_len; ++i) pts[0][i] = F(point[i]);
    std::vector<size_t> ch_off(layers), slot_off(layers), pt_off(layers), eq_off(layers);
    size_t pt_total = 0, eq_total = 0, scratch_total = 0;
    {
        size_t c = 0, slot = 0;
        for (int l = 0; l < layers; ++l) {
            const int n = (int)pts[l].size();
            ch_off[l] = c;
            slot_off[l] = slot;
            slot += (size_t)n;
            std::vector<f128> next(n + 1);
            for (int i = n - 1; i >= 0; --i) next[i] = F(challenges[c++]);
            next[n] = F(challenges[c++]);
            pts[l + 1] = next;
            pt_off[l] = pt_total;
            pt_total += (size_t)n;
            eq_off[l] = eq_total;
            eq_total += (size_t)1 << (n - 1);
            if (n - 1 > bitz::kGkrEqSmallMax)
                scratch_total += ((size_t)1 << ((n - 1) / 2)) + ((size_t)1 << (n - 1 - (n - 1) / 2));
        }
    }
    // Descriptors.
    std::vector<bitz::GkrEqItem> eq_items;
    std::vector<bitz::GkrCombine> combines;
    std::vector<uint32_t> comb_map;
    Dev pdev, chdev, rt4, eqa, eqs, desc;
    TRY(pdev.alloc(std::max<size_t>(pt_total, 1) * sizeof(f128)));
    TRY(chdev.alloc(std::max<size_t>(need, 1) * sizeof(f128)));
    TRY(rt4.alloc(std::max<size_t>(need, 1) * kConstTable4Entries * sizeof(uint4)));
    TRY(eqa.alloc(std::max<size_t>(eq_total, 1) * sizeof(f128)));
    TRY(eqs.alloc(std::max<size_t>(scratch_total, 1) * sizeof(f128)));
    TRY(h->acc.alloc((total_rounds + (size_t)layers) * 8 * sizeof(uint32_t)));
    {
        size_t so = 0;
        for (int l = 0; l < layers; ++l) {
            const int n = (int)pts[l].size(), k = n - 1;
            const f128* r = pdev.f() + pt_off[l];
            if (k <= bitz::kGkrEqSmallMax) {
                eq_items.push_back({r, k, eqa.f() + eq_off[l]});
                continue;
            }
            const int kl = k / 2, kh = k - kl;
            f128* low = eqs.f() + so;
            f128* high = low + ((size_t)1 << kl);
            so += ((size_t)1 << kl) + ((size_t)1 << kh);
            eq_items.push_back({r, kl, low});
            eq_items.push_back({r + kl, kh, high});
            bitz::GkrCombine c{};
            c.low = low;
            c.high = high;
            c.out = eqa.f() + eq_off[l];
            c.n = (size_t)1 << k;
            c.low_bits = kl;
            c.blocks = grid_for(c.n);
            c.block_base = (uint32_t)comb_map.size();
            comb_map.insert(comb_map.end(), c.blocks, (uint32_t)combines.size());
            combines.push_back(c);
        }
    }
    auto layer_ptrs = [&](int l, f128** left, f128** right) {
        const int j = layers - 1 - l;  // storage layer (0 = leaves)
        *left = h->store.f() + h->off[j];
        *right = *left + ((size_t)1 << pts[l].size());
    };
    int max_n = 0;
    for (int l = 0; l < layers; ++l) max_n = std::max(max_n, (int)pts[l].size());
    std::vector<std::vector<bitz::GkrRound>> rounds;
    std::vector<std::vector<uint32_t>> round_maps;
    for (int i = max_n - 1; i > T; --i) {
        std::vector<bitz::GkrRound> items;
        std::vector<uint32_t> map;
        for (int l = 0; l < layers; ++l) {
            const int n = (int)pts[l].size();
            if (i > n - 1) continue;
            bitz::GkrRound g{};
            layer_ptrs(l, &g.left, &g.right);
            g.eq = eqa.f() + eq_off[l];
            g.h = (size_t)1 << i;
            g.first = i == n - 1;
            // bind by r_{i+1}: the layer's challenge table of coordinate i+1
            if (!g.first) g.rt4 = static_cast<const uint4*>(rt4.p) + (ch_off[l] + (size_t)(n - 1 - (i + 1))) * kConstTable4Entries;
            g.at_one = hostf::is_zero(pts[l][i]);
            g.acc = h->acc.u32() + 8 * (slot_off[l] + (size_t)(n - 1 - i));
            g.blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>((g.h + bitz::kStageBlock - 1) / bitz::kStageBlock, 2048));
            g.block_base = (uint32_t)map.size();
            map.insert(map.end(), g.blocks, (uint32_t)items.size());
            items.push_back(g);
        }
        rounds.push_back(std::move(items));
        round_maps.push_back(std::move(map));
    }
    std::vector<bitz::GkrTail> tails(layers);
    uint32_t* evals = h->acc.u32() + 8 * total_rounds;
    for (int l = 0; l < layers; ++l) {
        const int n = (int)pts[l].size();
        auto& t = tails[l];
        layer_ptrs(l, &t.left, &t.right);
        t.eq = eqa.f() + eq_off[l];
        t.rtabs4 = static_cast<const uint4*>(rt4.p) + ch_off[l] * kConstTable4Entries;
        t.acc = h->acc.u32() + 8 * slot_off[l];
        t.evals = evals + 8 * (size_t)l;
        t.at_one = 0;
        for (int i = 0; i < n; ++i)
            if (hostf::is_zero(pts[l][i])) t.at_one |= 1u << i;
        t.n = n;
        t.i0 = std::min(n - 1, T);
    }
    // Staging: points | challenges | eq items | combines | map | rounds + maps | tails.
    size_t off = 0;
    auto place = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        const size_t o = off;
        off += bytes;
        return o;
    };
    const size_t eqi_o = place(eq_items.size() * sizeof(bitz::GkrEqItem));
    const size_t cmb_o = place(combines.size() * sizeof(bitz::GkrCombine));
    const size_t cmap_o = place(comb_map.size() * 4);
    std::vector<size_t> rd_o(rounds.size()), rm_o(rounds.size());
    for (size_t g = 0; g < rounds.size(); ++g) {
        rd_o[g] = place(rounds[g].size() * sizeof(bitz::GkrRound));
        rm_o[g] = place(round_maps[g].size() * 4);
    }
    const size_t tl_o = place(tails.size() * sizeof(bitz::GkrTail));
    const size_t desc_bytes = off;
    std::vector<char> stage(desc_bytes);
    memcpy(stage.data() + eqi_o, eq_items.data(), eq_items.size() * sizeof(bitz::GkrEqItem));
    memcpy(stage.data() + cmb_o, combines.data(), combines.size() * sizeof(bitz::GkrCombine));
    memcpy(stage.data() + cmap_o, comb_map.data(), comb_map.size() * 4);
    for (size_t g = 0; g < rounds.size(); ++g) {
        memcpy(stage.data() + rd_o[g], rounds[g].data(), rounds[g].size() * sizeof(bitz::GkrRound));
        memcpy(stage.data() + rm_o[g], round_maps[g].data(), round_maps[g].size() * 4);
    }
    memcpy(stage.data() + tl_o, tails.data(), tails.size() * sizeof(bitz::GkrTail));
    std::vector<f128> flat(std::max<size_t>(pt_total, 1));
    for (int l = 0; l < layers; ++l) std::copy(pts[l].begin(), pts[l].end(), flat.begin() + pt_off[l]);
    TRY(desc.alloc(std::max<size_t>(desc_bytes, 16)));
    HIP_TRY(hipMemcpyAsync(desc.p, stage.data(), desc_bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(pdev.p, flat.data(), flat.size() * sizeof(f128), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(chdev.p, challenges, need * sizeof(f128), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(h->acc.p, 0, h->acc.bytes, s));
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_tail),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kGkrTailSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_round_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kGkrRoundSmem);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_eq_items),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&bitz::k_gkr_eq_combine_batch),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bitz::kLaneSmem256);
        attr = true;
    }
    const char* dp = static_cast<const char*>(desc.p);
    hipLaunchKernelGGL(bitz::k_const_tables4, dim3((unsigned)need), dim3(512), 0, s, chdev.f(),
                       static_cast<uint4*>(rt4.p));
    TRY(check_launch());
    hipLaunchKernelGGL(bitz::k_gkr_eq_items, dim3((unsigned)eq_items.size()), dim3(bitz::kStageBlock),
                       bitz::kLaneSmem256, s,
                       reinterpret_cast<const bitz::GkrEqItem*>(dp + eqi_o));
    TRY(check_launch());
    if (!combines.empty()) {
        hipLaunchKernelGGL(bitz::k_gkr_eq_combine_batch, dim3((unsigned)comb_map.size()), dim3(bitz::kStageBlock),
                           bitz::kLaneSmem256, s,
                           reinterpret_cast<const bitz::GkrCombine*>(dp + cmb_o),
                           reinterpret_cast<const uint32_t*>(dp + cmap_o));
        TRY(check_launch());
    }
    // The round launches run back to back: one event pair spans them.
    std::vector<Timed> ts(2);
    double round_bytes = 0;
    for (const auto& rg : rounds)
        for (const auto& r : rg) round_bytes += (r.first ? 80.0 : 240.0) * (double)r.h;
    for (size_t g = 0; g < rounds.size(); ++g) {
        if (g == 0) {
            ts[0].begin(s, "gkr_round_batch", round_bytes);
            ts[0].launches = rounds.size();
        }
        hipLaunchKernelGGL(bitz::k_gkr_round_batch, dim3((unsigned)round_maps[g].size()), dim3(bitz::kStageBlock),
                           bitz::kGkrRoundSmem, s,
                           reinterpret_cast<const bitz::GkrRound*>(dp + rd_o[g]),
                           reinterpret_cast<const uint32_t*>(dp + rm_o[g]));
        if (g + 1 == rounds.size()) ts[0].end(s);
        TRY(check_launch());
    }
    ts.back().begin(s, "gkr_tail", 0.0);
    hipLaunchKernelGGL(bitz::k_gkr_tail, dim3((unsigned)layers), dim3(bitz::kStageBlock), bitz::kGkrTailSmem, s,
                       reinterpret_cast<const bitz::GkrTail*>(dp + tl_o));
    ts.back().end(s);
    TRY(check_launch());
    Pinned down;
    const size_t words = 8 * (total_rounds + (size_t)layers);
    TRY(down.alloc(std::max<size_t>(words, 1) * 4));
    HIP_TRY(hipMemcpyAsync(down.p, h->acc.p, words * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (auto& t : ts) t.collect();
    const uint32_t* w = static_cast<const uint32_t*>(down.p);
    size_t slot = 0;
    f128 claim{0, 0};
    for (int l = 0; l < layers; ++l) {
        const int n = (int)pts[l].size();
        f128 eq_bound = hostf::one();
        for (int i = n - 1; i >= 0; --i) {
            f128 sums[2];
            memcpy(sums, w + 8 * slot, 32);
            // lib.rs:142: prover_message([eq_bound * sum_b, eq_bound * sum_inf]).
            messages[2 * slot] = B(hostf::mul(eq_bound, sums[0]));
            messages[2 * slot + 1] = B(hostf::mul(eq_bound, sums[1]));
            const f128 r = pts[l + 1][i], z = pts[l][i];
            eq_bound = hostf::mul(eq_bound, f128{1 ^ r.lo ^ z.lo, r.hi ^ z.hi});  // eq_factor
            ++slot;
        }
        f128 lr[2];
        memcpy(lr, w + 8 * total_rounds + 8 * l, 32);
        layer_evals[2 * l] = B(lr[0]);
        layer_evals[2 * l + 1] = B(lr[1]);
        const f128 tt = pts[l + 1][n];
        claim = hostf::add(lr[0], hostf::mul(tt, hostf::add(lr[1], lr[0])));
    }
    for (size_t i = 0; i < pts[layers].size(); ++i) leaf_point[i] = B(pts[layers][i]);
    *leaf_claim = B(claim);
    return BITZ_OK;
}

extern "C" void bitz_gkr_free(bitz_gkr* h) {
    if (h) {
        (void)hipStreamSynchronize(h->so.s);
        delete h;
    }
}

// ============================================================ serving =====
// A proving server for SHA-256 compression-chain requests, native end to end:
// admission in arrival order by slots and HBM budget, preparer threads
// (witness on the host, then an engine proof from the packed assignment),
// background PIOP engines, and harvester threads that take finished proofs
// and free their slot.  No reference counterpart (the reference proves one
// statement per process, tooling/cli/src/cmd/circuit_e2e.rs:48-77).
namespace {
struct ServeCircuit {
    const bitz_sha256* c = nullptr;
    bitz_matrices* m = nullptr;
    uint8_t digest[32] = {};
    uint64_t footprint = 0;
    uint64_t resident = 0;  // its matrices and the engines' bind buffers for it
    size_t columns = 0, witness_bits = 0;
};
}  // namespace

struct bitz_server {
    uint32_t slots = 8, preparers = 4;
    uint64_t budget = 0;
    // Resident for the server's life: every circuit's matrices and each
    // engine's bind buffers for it; the requests share the rest of the budget.
    uint64_t resident = 0;
    std::vector<bitz_piop_engine*> engines;
    std::vector<ServeCircuit> circuits;
    ~bitz_server() {
        for (auto* e : engines) bitz_piop_engine_free(e);
    }
};

extern "C" int32_t bitz_server_create(bitz_server** out, uint32_t slots, uint32_t engines, uint32_t preparers,
                                      uint64_t hbm_budget) {
    if (!out) return fail(BITZ_ERR_ARG, "null pointer");
    if (!slots || !engines || !preparers) return fail(BITZ_ERR_ARG, "slots, engines and preparers must be positive");
    auto* s = new bitz_server;
    s->slots = slots;
    s->preparers = preparers;
    s->budget = hbm_budget;
    for (uint32_t k = 0; k < engines; ++k) {
        bitz_piop_engine* e = nullptr;
        int32_t rc = bitz_piop_engine_create(&e);
        if (!rc) rc = bitz_piop_engine_run(e);
        if (rc) {
            if (e) bitz_piop_engine_free(e);
            delete s;
            return rc;
        }
        s->engines.push_back(e);
    }
    *out = s;
    return BITZ_OK;
}

extern "C" int32_t bitz_server_add_circuit(bitz_server* s, const bitz_sha256* c, bitz_matrices* m,
                                           const uint8_t digest[32], uint64_t footprint, uint32_t* id) {
    if (!s || !c || !m || !digest || !id) return fail(BITZ_ERR_ARG, "null pointer");
    ServeCircuit sc;
    sc.c = c;
    sc.m = m;
    memcpy(sc.digest, digest, 32);
    sc.footprint = footprint;
    size_t rows = 0, mpos = 0, nnz[3];
    TRY(bitz_sha256_shape(c, &rows, &sc.columns, &sc.witness_bits, nnz, &mpos));
    // The matrices, and per engine the bind buffers (BindScratch): the point,
    // the row eq table, three row scales and one product per nonzero.
    const uint64_t bind = 16ull * ((uint64_t)std::max(m->row_vars, 1) + 4ull * m->rows_padded + m->nnz);
    sc.resident = matrices_device_bytes(m) + (uint64_t)s->engines.size() * bind;
    s->resident += sc.resident;
    *id = (uint32_t)s->circuits.size();
    s->circuits.push_back(sc);
    return BITZ_OK;
}

extern "C" int32_t bitz_server_resident_bytes(const bitz_server* s, uint64_t* bytes) {
    if (!s || !bytes) return fail(BITZ_ERR_ARG, "null pointer");
    *bytes = s->resident;
    return BITZ_OK;
}

extern "C" int32_t bitz_server_serve(bitz_server* s, bitz_serve_request* reqs, size_t count) {
    if (!s || (count && !reqs)) return fail(BITZ_ERR_ARG, "null pointer");
    for (size_t i = 0; i < count; ++i)
        if (reqs[i].circuit >= s->circuits.size() || !reqs[i].message)
            return fail(BITZ_ERR_ARG, "request " + std::to_string(i) + ": no such circuit or no message");
    std::vector<size_t> order(count);
    for (size_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    std::mutex mu;
    std::condition_variable cv;
    std::deque<size_t> ready;  // admitted, waiting for a preparer
    std::vector<uint32_t> load(s->engines.size(), 0);
    std::vector<std::unique_ptr<bitz_transcript>> ts(count);
    size_t left = count, in_flight = 0;
    uint64_t used = 0;
    int32_t err = BITZ_OK;
    std::string err_msg;
    bool stop = false;
    const auto t0 = std::chrono::steady_clock::now();
    auto now_s = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    auto fail_all = [&](int32_t rc) {  // under mu
        if (!err) {
            err = rc;
            err_msg = bitz_last_error();
        }
        cv.notify_all();
    };
    // Harvester threads wait on their engine's condition; quit wakes them.
    std::atomic<bool> quit{false};
    auto wake_harvesters = [&] {
        quit = true;
        for (auto* e : s->engines) {
            { std::lock_guard<std::mutex> lock(e->mu); }
            e->cv.notify_all();
        }
    };
    // Preparers: the witness (host), then the proof joins the least loaded engine.
    auto prepare = [&] {
        std::vector<uint64_t> f, h;
        for (;;) {
            size_t i;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv.wait(lock, [&] { return stop || !ready.empty() || err; });
                if (stop || err) return;
                i = ready.front();
                ready.pop_front();
            }
            bitz_serve_request& r = reqs[i];
            const ServeCircuit& c = s->circuits[r.circuit];
            f.resize(std::max<size_t>((c.witness_bits + 63) / 64, 1));
            h.resize(std::max<size_t>((c.columns + 63) / 64, 1));
            int32_t sat = 0;
            int32_t rc = bitz_sha256_witness(c.c, r.message, r.iv, f.data(), h.data(), r.digest, &sat);
            const std::string dom = "bitz/serve/" + std::to_string(r.rid);
            ts[i].reset(bitz_transcript_keccak(reinterpret_cast<const uint8_t*>(dom.data()), dom.size()));
            size_t k = 0;
            {
                std::lock_guard<std::mutex> lock(mu);
                for (size_t e = 1; e < load.size(); ++e)
                    if (load[e] < load[k]) k = e;
                load[k] += 1;
                r.engine = (uint32_t)k;
            }
            if (!rc && !ts[i]) rc = fail(BITZ_ERR_ALLOC, "transcript creation failed");
            if (!rc)
                rc = bitz_piop_engine_submit_assignment(s->engines[k], i, ts[i].get(), c.m, c.digest, h.data(),
                                                        c.columns, c.columns);
            if (rc) {
                {
                    std::lock_guard<std::mutex> lock(mu);
                    r.status = rc;
                    fail_all(rc);
                }
                wake_harvesters();
                return;
            }
        }
    };
    // Harvesters: one per engine, take each finished proof, free its slot.
    auto harvest = [&](size_t k) {
        bitz_piop_engine* e = s->engines[k];
        std::vector<uint64_t> tags(64);
        for (;;) {
            bool any = false;
            {
                std::unique_lock<std::mutex> lock(e->mu);
                e->cv.wait_for(lock, std::chrono::milliseconds(20),
                               [&] { return !e->finished.empty() || e->worker_rc || quit.load(); });
                any = !e->finished.empty() || e->worker_rc;
                if (!any && quit.load()) return;
            }
            if (!any) continue;
            size_t n = 0;
            int32_t rc = bitz_piop_engine_done(e, tags.data(), tags.size(), &n);
            for (size_t q = 0; !rc && q < std::min(n, tags.size()); ++q) {
                bitz_serve_request& r = reqs[tags[q]];
                rc = bitz_piop_engine_take(e, tags[q], r.outer_rounds, r.outer_claims, r.inner_rounds, r.inner_point,
                                           r.matrix_eval, r.final_claim);
                r.status = rc;
                r.done_s = now_s();
                std::lock_guard<std::mutex> lock(mu);
                used -= s->circuits[r.circuit].footprint;
                in_flight -= 1;
                load[k] -= 1;
                left -= 1;
                cv.notify_all();
            }
            if (rc) {
                {
                    std::lock_guard<std::mutex> lock(mu);
                    fail_all(rc);
                }
                wake_harvesters();
                return;
            }
        }
    };
    std::vector<std::thread> threads;
    for (uint32_t p = 0; p < s->preparers; ++p) threads.emplace_back(prepare);
    for (size_t k = 0; k < s->engines.size(); ++k) threads.emplace_back(harvest, k);
    {
        // Admission in arrival order: a slot free and the footprint within
        // what the resident circuits leave of the budget (a request larger
        // than that runs alone).
        std::unique_lock<std::mutex> lock(mu);
        const uint64_t budget = s->budget > s->resident ? s->budget - s->resident : 0;
        size_t next = 0;
        while (left && !err) {
            const double t = now_s();
            while (next < count && reqs[order[next]].arrival_s <= t && in_flight < s->slots) {
                bitz_serve_request& r = reqs[order[next]];
                const uint64_t fp = s->circuits[r.circuit].footprint;
                if (in_flight && used + fp > budget) break;
                r.admitted_s = t;
                used += fp;
                in_flight += 1;
                ready.push_back(order[next]);
                ++next;
                cv.notify_all();
            }
            if (next < count && in_flight < s->slots && (!in_flight || used + s->circuits[reqs[order[next]].circuit].footprint <= budget)) {
                const double wait = reqs[order[next]].arrival_s - now_s();
                if (wait > 0) cv.wait_for(lock, std::chrono::duration<double>(wait));
            } else {
                cv.wait(lock);
            }
        }
        stop = true;
        cv.notify_all();
    }
    wake_harvesters();
    for (auto& th : threads) th.join();
    if (err) {
        // Leave the engines clean: wait for what is in flight and drop it.
        for (auto* e : s->engines) {
            size_t waiting = 0, n = 0;
            std::vector<uint64_t> tags(count ? count : 1);
            for (int tries = 0; tries < 1000; ++tries) {
                (void)bitz_piop_engine_wait(e, 1000, &waiting);
                if (bitz_piop_engine_done(e, tags.data(), tags.size(), &n) == BITZ_OK)
                    for (size_t q = 0; q < std::min(n, tags.size()); ++q) {
                        bitz_serve_request& r = reqs[tags[q]];
                        (void)bitz_piop_engine_take(e, tags[q], r.outer_rounds, r.outer_claims, r.inner_rounds,
                                                    r.inner_point, r.matrix_eval, r.final_claim);
                    }
                size_t in = 0;
                {
                    std::lock_guard<std::mutex> lock(e->mu);
                    in = e->jobs.size() + e->incoming.size() + e->finished.size();
                }
                if (!in) break;
            }
        }
        return fail(err, err_msg);
    }
    return BITZ_OK;
}

extern "C" void bitz_server_free(bitz_server* s) { delete s; }
