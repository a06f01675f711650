// C-ABI entry points of libseqtree_hip.so (declared in include/seqtree_hip.h).
//
// Each host-memory entry point copies the CSR inputs to the device, runs the same device
// path as the *_device entry point on a private stream, and copies the matrix back. Scratch
// is stream-ordered (hipMallocAsync / hipFreeAsync on the call's stream), so concurrent
// calls never share buffers. Integer-width choices are made here on the host from the input
// bounds (a range proof), never by trying and retrying on the device.

#include "gapblock_core.hpp"
#include "st_internal.hpp"
#include "../../include/seqtree_hip.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

// Device scratch of st_gapblock_matrix_device_noalloc: one array per GPU (each device holds
// its own copy at its own address, kept until the process ends). Host code never uses this
// name as an address -- it is a host-side stand-in -- but the device address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch below).
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

using namespace st;

constexpr const char* kVersion = "seqtree_hip 0.1.0 (gfx950)";

void set_status(st_status* st, int code, const char* fmt, ...) {
    if (!st) return;
    st->code = code;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(st->message, sizeof(st->message), fmt, ap);
    va_end(ap);
}

int ok(st_status* st) {
    if (st) {
        st->code = ST_OK;
        st->message[0] = 0;
    }
    return ST_OK;
}

struct HipError {
    int code;
    std::string msg;
};

void check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                         ? ST_OUT_OF_MEMORY
                         : ST_DEVICE_ERROR;
    throw HipError{code, std::string(what) + ": " + hipGetErrorString(e)};
}

// Checks once per process that device 0 is reachable; the code object is gfx950 only.
void require_device() {
    static std::once_flag once;
    static int state = 0;  // 1 ok, -1 no device, -2 wrong arch
    static std::string arch;
    std::call_once(once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            state = -1;
            return;
        }
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
            state = -1;
            return;
        }
        arch = p.gcnArchName;
        state = arch.rfind("gfx950", 0) == 0 ? 1 : -2;
    });
    if (state == -1) throw HipError{ST_DEVICE_ERROR, "seqtree_hip: no HIP device is available"};
    if (state == -2)
        throw HipError{ST_DEVICE_ERROR,
                       "seqtree_hip: device is " + arch + "; this library is built for gfx950"};
}

bool pools_supported() {
    static std::once_flag once;
    static bool yes = false;
    std::call_once(once, [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMemoryPoolsSupported, dev) == hipSuccess)
            yes = v != 0;
    });
    return yes;
}

// Freed stream-ordered memory stays in the current device's default pool up to this many
// bytes instead of going back to the driver at the next synchronisation (the pool's default
// threshold is zero), so a call's scratch is mapped once and reused by the calls after it.
constexpr uint64_t kPoolKeepBytes = uint64_t(4) << 30;

void keep_pool_memory() {
    constexpr int kMaxDevices = 64;
    static std::once_flag once[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return;
    std::call_once(once[dev], [dev] {
        hipMemPool_t pool = nullptr;
        uint64_t keep = kPoolKeepBytes;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool)
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    });
}

// Stream-ordered scratch for one call. Everything allocated here is released on the same
// stream when the arena goes out of scope (after a stream sync when pools are unavailable).
// Requests of at most kSmall bytes are carved (256-byte aligned) out of shared blocks of
// kBlock bytes, larger ones get their own allocation: a stream-ordered free costs the host
// ~10-20 us on this runtime, so a call with a handful of small buffers frees once.
//
// A fixed arena carves every request from one caller-given device region instead, allocating
// nothing, and throws ST_INVALID_ARGUMENT when the region is exhausted.
class Arena {
public:
    explicit Arena(hipStream_t s) : s_(s) {}
    Arena(hipStream_t s, void* region, size_t bytes)
        : s_(s), block_(static_cast<char*>(region)), fixed_bytes_(region ? bytes : 0) {}
    bool fixed() const { return fixed_bytes_ != 0; }
    // A device region (owned by the caller, reused stream-ordered on s) that requests use
    // first, falling back to allocations once it is full.
    void prefer(void* region, size_t bytes) {
        pref_ = static_cast<char*>(region);
        pref_bytes_ = bytes;
    }
    ~Arena() {
        if (bufs_.empty()) return;
        if (pools_supported()) {
            for (void* p : bufs_) (void)hipFreeAsync(p, s_);
        } else {
            (void)hipStreamSynchronize(s_);
            for (void* p : bufs_) (void)hipFree(p);
        }
    }
    template <class T>
    T* get(size_t count) {
        const size_t bytes = (std::max<size_t>(count * sizeof(T), 1) + 255) / 256 * 256;
        if (fixed()) {
            if (used_ + bytes > fixed_bytes_)
                throw HipError{ST_INVALID_ARGUMENT,
                               "seqtree_hip: the call needs more than the " +
                                   std::to_string(fixed_bytes_ >> 20) +
                                   " MiB of static scratch"};
            T* p = reinterpret_cast<T*>(block_ + used_);
            used_ += bytes;
            return p;
        }
        if (pref_ && pref_used_ + bytes <= pref_bytes_) {
            T* p = reinterpret_cast<T*>(pref_ + pref_used_);
            pref_used_ += bytes;
            return p;
        }
        if (bytes > kSmall) return static_cast<T*>(raw(bytes));
        if (block_ == nullptr || used_ + bytes > kBlock) {
            block_ = static_cast<char*>(raw(kBlock));
            used_ = 0;
        }
        T* p = reinterpret_cast<T*>(block_ + used_);
        used_ += bytes;
        return p;
    }

private:
    static constexpr size_t kSmall = size_t(256) << 10;
    static constexpr size_t kBlock = size_t(1) << 20;

    void* raw(size_t bytes) {
        void* p = nullptr;
        if (pools_supported()) {
            keep_pool_memory();
            check(hipMallocAsync(&p, bytes, s_), "hipMallocAsync");
        } else {
            check(hipMalloc(&p, bytes), "hipMalloc");
        }
        bufs_.push_back(p);
        return p;
    }

    hipStream_t s_;
    std::vector<void*> bufs_;
    char* block_ = nullptr;
    size_t used_ = 0;
    size_t fixed_bytes_ = 0;
    char* pref_ = nullptr;
    size_t pref_bytes_ = 0, pref_used_ = 0;
};

template <class T>
void h2d(T* dst, const T* src, size_t count, hipStream_t s) {
    if (count) check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s), "H2D");
}

template <class T>
void d2h(T* dst, const T* src, size_t count, hipStream_t s) {
    if (count) check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s), "D2H");
}

void sync(hipStream_t s) { check(hipStreamSynchronize(s), "hipStreamSynchronize"); }

// ---- host <-> device transfers of the host-memory entry points ----------------------------
// Every host buffer a host-memory entry point reads or writes goes through pinned staging owned
// by the library: the runtime's own path for pageable memory is slower, and where it pins the
// caller's pages in place (large page-aligned buffers) the platform may refuse it. Copies
// between the pinned slots and the caller's memory are split over a small pool of threads.

// Persistent worker threads for host memcpy: copy() splits one copy into equal parts, the
// calling thread takes the first, and returns when every part is done. One copy at a time.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool* pool = new CopyPool();  // never destroyed: workers live with the process
        return *pool;
    }
    void copy(void* dst, const void* src, size_t bytes) {
        const size_t parts = bytes < (size_t(1) << 20) ? 1 : std::min<size_t>(workers_ + 1,
                                                                               bytes >> 19);
        if (parts <= 1) {
            std::memcpy(dst, src, bytes);
            return;
        }
        std::lock_guard<std::mutex> one(copy_mu_);
        const size_t step = (bytes / parts + 63) & ~size_t(63);
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            step_ = step;
            next_ = 1;
            parts_ = parts;
            pending_ = parts - 1;
            ++generation_;
        }
        cv_.notify_all();
        std::memcpy(dst, src, std::min(step, bytes));
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        workers_ = std::min(7u, hw > 1 ? hw - 1 : 0u);
        for (unsigned w = 0; w < workers_; ++w) std::thread([this] { run(); }).detach();
    }
    void run() {
        uint64_t seen = 0;
        for (;;) {
            size_t part;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return generation_ != seen && next_ < parts_; });
                part = next_++;
                if (next_ >= parts_) seen = generation_;
            }
            const size_t at = part * step_;
            if (at < bytes_) std::memcpy(dst_ + at, src_ + at, std::min(step_, bytes_ - at));
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_all();
        }
    }
    unsigned workers_ = 0;
    std::mutex copy_mu_, mu_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0, step_ = 0, next_ = 0, parts_ = 0, pending_ = 0;
    uint64_t generation_ = 0;
};

// Host buffers pinned by st_host_alloc, by base address. A download whose destination lies in
// one goes straight into it (one DMA, no staging and no host copy).
std::mutex g_pin_mu;
std::map<uintptr_t, size_t> g_pinned;

bool library_pinned(const void* p, size_t bytes) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pinned.upper_bound(a);
    if (it == g_pinned.begin()) return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}

// One transfer context: a stream, two pinned staging slots with an event each, and a small
// device block for the scratch of small calls. A host-memory call holds one for its duration
// (HostCall); the pool keeps them for reuse, so it grows to the number of concurrent calls.
constexpr size_t kStageBytes = size_t(16) << 20;
constexpr size_t kContextScratch = size_t(4) << 20;

struct HostContext {
    int device = 0;
    hipStream_t stream = nullptr;  // uploads and scoring
    hipStream_t copy = nullptr;    // downloads that overlap later scoring (download_after)
    uint8_t* slot[2] = {nullptr, nullptr};
    hipEvent_t ready[2] = {nullptr, nullptr};
    void* scratch = nullptr;
};

std::mutex g_ctx_mu;
std::vector<HostContext*> g_ctx_free;

class HostCall {
public:
    HostCall() {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        {
            std::lock_guard<std::mutex> lk(g_ctx_mu);
            for (size_t i = 0; i < g_ctx_free.size(); ++i)
                if (g_ctx_free[i]->device == dev) {
                    c_ = g_ctx_free[i];
                    g_ctx_free.erase(g_ctx_free.begin() + i);
                    break;
                }
        }
        if (!c_) c_ = create(dev);
    }
    ~HostCall() {
        // the streams are drained before the context (and its slots) can serve another call
        if (hipStreamSynchronize(c_->stream) != hipSuccess ||
            hipStreamSynchronize(c_->copy) != hipSuccess)
            return;  // leak a broken context
        release_marks();
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        g_ctx_free.push_back(c_);
    }
    HostCall(const HostCall&) = delete;
    HostCall& operator=(const HostCall&) = delete;

    hipStream_t stream() const { return c_->stream; }
    void* scratch() const { return c_->scratch; }

    // host -> device, stream-ordered: each chunk is copied into a slot once that slot's
    // previous transfer has finished, then sent.
    void upload(void* dev, const void* host, size_t bytes) {
        const char* src = static_cast<const char*>(host);
        char* dst = static_cast<char*>(dev);
        for (size_t at = 0; at < bytes; at += kStageBytes) {
            const size_t n = std::min(kStageBytes, bytes - at);
            const int k = turn_++ & 1;
            check(hipEventSynchronize(c_->ready[k]), "hipEventSynchronize");
            CopyPool::get().copy(c_->slot[k], src + at, n);
            check(hipMemcpyAsync(dst + at, c_->slot[k], n, hipMemcpyHostToDevice, c_->stream),
                  "H2D");
            check(hipEventRecord(c_->ready[k], c_->stream), "hipEventRecord");
        }
    }
    template <class T>
    void put(T* dev, const T* host, size_t count) { upload(dev, host, count * sizeof(T)); }

    // device -> host after everything enqueued so far on the call's stream: chunk c + 1
    // travels while chunk c is copied out of its slot. Returns when the data is in `host`.
    void download(void* host, const void* dev, size_t bytes) {
        if (download_on(c_->stream, host, dev, bytes)) sync(c_->stream);
    }
    // The same on the copy stream once `after` (recorded on the call's stream after the
    // uploads) has completed, so the scoring enqueued after `after` overlaps the transfer.
    // Into st_host_alloc memory the transfer is only enqueued: finish() waits for it.
    void download_after(hipEvent_t after, void* host, const void* dev, size_t bytes) {
        check(hipStreamWaitEvent(c_->copy, after, 0), "hipStreamWaitEvent");
        download_on(c_->copy, host, dev, bytes);
    }
    // Waits for every transfer download_after has enqueued.
    void finish() { sync(c_->copy); }
    // An event recorded on the call's stream now (one of a reusable set).
    hipEvent_t mark() {
        if (marks_used_ == c_marks_.size()) {
            hipEvent_t e;
            check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            c_marks_.push_back(e);
        }
        hipEvent_t e = c_marks_[marks_used_++];
        check(hipEventRecord(e, c_->stream), "hipEventRecord");
        return e;
    }
    template <class T>
    void get(T* host, const T* dev, size_t count) { download(host, dev, count * sizeof(T)); }

private:
    // Returns true when the transfer went straight into library-pinned memory and is only
    // enqueued on `st`; otherwise the data is in `host` on return.
    bool download_on(hipStream_t st, void* host, const void* dev, size_t bytes) {
        if (bytes == 0) return false;
        if (library_pinned(host, bytes)) {
            check(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st), "D2H");
            return true;
        }
        const char* src = static_cast<const char*>(dev);
        char* dst = static_cast<char*>(host);
        const size_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
        auto send = [&](size_t c) {
            const int k = int((turn_ + c) & 1);
            const size_t at = c * kStageBytes, n = std::min(kStageBytes, bytes - at);
            check(hipMemcpyAsync(c_->slot[k], src + at, n, hipMemcpyDeviceToHost, st), "D2H");
            check(hipEventRecord(c_->ready[k], st), "hipEventRecord");
        };
        // both slots' earlier uploads have been read once the stream reaches this point
        send(0);
        for (size_t c = 0; c < chunks; ++c) {
            if (c + 1 < chunks) send(c + 1);
            const int k = int((turn_ + c) & 1);
            check(hipEventSynchronize(c_->ready[k]), "hipEventSynchronize");
            const size_t at = c * kStageBytes, n = std::min(kStageBytes, bytes - at);
            CopyPool::get().copy(dst + at, c_->slot[k], n);
        }
        turn_ += chunks;
        return false;
    }
    static HostContext* create(int dev) {
        auto* c = new HostContext();
        c->device = dev;
        check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
        check(hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking), "stream");
        for (int k = 0; k < 2; ++k) {
            check(hipHostMalloc(reinterpret_cast<void**>(&c->slot[k]), kStageBytes,
                                hipHostMallocDefault),
                  "hipHostMalloc");
            check(hipEventCreateWithFlags(&c->ready[k], hipEventDisableTiming), "event");
            check(hipEventRecord(c->ready[k], c->stream), "hipEventRecord");
        }
        check(hipMalloc(&c->scratch, kContextScratch), "hipMalloc");
        return c;
    }
    HostContext* c_ = nullptr;
    uint64_t turn_ = 0;
    std::vector<hipEvent_t> c_marks_;
    size_t marks_used_ = 0;

public:
    // events of mark() are released with the call (after the streams have drained)
    void release_marks() {
        for (hipEvent_t e : c_marks_) (void)hipEventDestroy(e);
        c_marks_.clear();
        marks_used_ = 0;
    }
};

// ---- optional launch timing (st_timing_enable / st_timing_collect) ----------------------
// When enabled, every kernel launch below is bracketed by two hipEvents recorded on the
// stream the kernel runs on; st_timing_collect resolves them into per-label totals.
struct Pending {
    std::string label;
    hipEvent_t start, stop;
};
std::mutex g_timing_mu;
bool g_timing = false;
std::vector<Pending> g_pending;
// set by the allocation-free entry point on its thread: no events are created meanwhile
thread_local bool t_no_alloc = false;

class LaunchTimer {
public:
    LaunchTimer(std::string label, hipStream_t s) : s_(s) {
        if (t_no_alloc) return;
        std::lock_guard<std::mutex> lk(g_timing_mu);
        if (!g_timing) return;
        on_ = hipEventCreate(&a_) == hipSuccess && hipEventCreate(&b_) == hipSuccess &&
              hipEventRecord(a_, s_) == hipSuccess;
        label_ = std::move(label);
    }
    ~LaunchTimer() {
        if (!on_) return;
        (void)hipEventRecord(b_, s_);
        std::lock_guard<std::mutex> lk(g_timing_mu);
        g_pending.push_back({label_, a_, b_});
    }

private:
    hipStream_t s_;
    bool on_ = false;
    hipEvent_t a_{}, b_{};
    std::string label_;
};

void launched(const char* what) { check(hipGetLastError(), what); }

uint32_t words_for(uint32_t max_len) { return (max_len + 3) / 4; }

// The packed layout's geometry for n references of at most max_len residues.
PackedRefs packed_geometry(uint64_t n, uint32_t max_len) {
    PackedRefs p{};
    p.nchunks = uint32_t((n + kChunk - 1) / kChunk);
    p.npad = p.nchunks * kChunk;
    p.nwords = words_for(max_len);
    return p;
}

// Packs the reference side (see st_internal.hpp for the layout); scale as launch_pack_refs.
// A twin job, when given, has its twin list built by the same launch.
PackedRefs pack_refs(Arena& ar, const uint8_t* bytes, const uint64_t* off, uint64_t n,
                     uint32_t max_len, hipStream_t s, uint32_t scale = 1,
                     const TwinJob& tw = TwinJob{}) {
    LaunchTimer t("pack_refs", s);
    PackedRefs p = packed_geometry(n, max_len);
    uint32_t* words = ar.get<uint32_t>(size_t(p.nwords + 1) * p.npad);
    uint32_t* len = ar.get<uint32_t>(p.npad);
    uint32_t* orig = ar.get<uint32_t>(p.npad);
    launch_pack_refs(bytes, off, n, p.nwords, p.npad, words, len, orig, scale, max_len, tw, s);
    launched("pack_refs");
    p.words = words;
    p.len = len;
    p.orig = orig;
    return p;
}

// Buckets queries by length; returns host copies of the class counts (one sync).
std::vector<uint32_t> classify(Arena& ar, const uint64_t* off, uint64_t n, const ClassBounds& b,
                               uint32_t** lists, hipStream_t s) {
    uint32_t* counts = ar.get<uint32_t>(kMaxClasses);
    *lists = ar.get<uint32_t>(size_t(b.count) * n);
    launch_fill_u32(counts, kMaxClasses, 0, s);
    launch_classify(off, n, b, counts, *lists, s);
    launched("classify");
    std::vector<uint32_t> h(kMaxClasses);
    d2h(h.data(), counts, kMaxClasses, s);
    sync(s);
    return h;
}

int64_t iabs64(int64_t v) { return v < 0 ? -v : v; }

// ---- device copies of small host tables --------------------------------------------------
// Penalty tables and prior cubes are a few KB to about a MB and usually the same from call to
// call; uploading them from pageable host memory on every call costs more GPU time than a
// small score matrix takes to compute. They are kept by content (device, exact values, zero
// padding) for the life of the process, read-only, so any number of calls on any streams can
// share one. At most kTableCacheMax entries / kTableCacheBytes; past that a call uploads its
// own stream-ordered copy as before.
struct DeviceTable {
    const int32_t* dev;
    int64_t abs_max;  // max |entry|
};
struct TableEntry {
    int device;
    std::vector<int32_t> host;
    size_t pad;
    int32_t* dev;
    int64_t abs_max;
};
constexpr size_t kTableCacheMax = 64;
constexpr size_t kTableCacheBytes = size_t(64) << 20;
std::mutex g_table_mu;
std::vector<TableEntry> g_tables;
size_t g_table_bytes = 0;

DeviceTable table_on_device(Arena& ar, const int32_t* host, size_t count, size_t pad,
                            hipStream_t s) {
    if (ar.fixed()) {  // no cache (its misses allocate): a stream-ordered copy into the arena
        int64_t abs_max = 0;
        for (size_t i = 0; i < count; ++i) abs_max = std::max(abs_max, iabs64(host[i]));
        int32_t* d = ar.get<int32_t>(count + pad);
        h2d(d, host, count, s);
        if (pad) launch_fill_u32(reinterpret_cast<uint32_t*>(d + count), pad, 0u, s);
        return {d, abs_max};
    }
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        for (const TableEntry& e : g_tables)
            if (e.device == dev && e.pad == pad && e.host.size() == count &&
                std::equal(host, host + count, e.host.begin()))
                return {e.dev, e.abs_max};
    }
    int64_t abs_max = 0;
    for (size_t i = 0; i < count; ++i) abs_max = std::max(abs_max, iabs64(host[i]));
    const size_t bytes = (count + pad) * sizeof(int32_t);
    {
        std::lock_guard<std::mutex> lk(g_table_mu);
        if (g_tables.size() < kTableCacheMax && g_table_bytes + bytes <= kTableCacheBytes) {
            std::vector<int32_t> staged(count + pad, 0);
            std::copy(host, host + count, staged.begin());
            int32_t* d = nullptr;
            check(hipMalloc(&d, bytes), "hipMalloc");
            if (hipMemcpy(d, staged.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(d);
                check(hipGetLastError(), "table upload");
            }
            g_tables.push_back({dev, std::vector<int32_t>(host, host + count), pad, d, abs_max});
            g_table_bytes += bytes;
            return {d, abs_max};
        }
    }
    int32_t* d = ar.get<int32_t>(count + pad);
    h2d(d, host, count, s);
    if (pad) launch_fill_u32(reinterpret_cast<uint32_t*>(d + count), pad, 0u, s);
    return {d, abs_max};
}

// ------------------------------------------------------------------------------------------
// gap-block
// ------------------------------------------------------------------------------------------
// Uploads the tables, decides the evaluation variant and (for the matrix and threshold modes)
// packs the references. Returns the launch arguments; `exact` / `profile` select the variant.
struct GapblockPlan {
    GapblockArgs a;
    bool exact;
    bool profile;
    bool short_path;
    bool twin;            // short launch with two queries of one length per workgroup
    uint2* twins = nullptr;  // its twin list (built by the packing launch)
};

// Twin rows (two queries in the 16-bit halves, gapblock_core.hpp) need every sum of a short
// launch inside int16: 6 T + R < 2^15 with T = q_max * max|pen| and R = max|prior|. Sets the
// 16-bit sentinel. (The pairing kernel indexes queries in 32 bits: N < 2^31.)
bool twin_launch(uint64_t N, uint32_t q_max, int64_t pen_abs, int64_t prior_abs, int32_t* big) {
    const int64_t T = int64_t(q_max) * pen_abs;
    if (N < 2 || N >= (uint64_t(1) << 31) || q_max > kTwinMaxLen || 6 * T + prior_abs >= 32767)
        return false;
    *big = int32_t(32767 - 2 * T);
    return true;
}

// Reference chunks per gap-block workgroup: at most what keeps a group's packed words around
// 512 KB (the workgroups resident together stream one group, see gapblock_kernel), at most
// what leaves about four workgroups per CU in all, and balanced so the groups are equal.
uint32_t chunk_group(uint64_t n_queries, uint32_t nchunks, uint32_t nwords) {
    if (nchunks <= 1) return 1;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint64_t units = n_queries * uint64_t(nchunks);
    const uint64_t by_fill = std::max<uint64_t>(1, units / (uint64_t(cus) * 4));
    const uint64_t chunk_bytes = uint64_t(kChunk) * (nwords + 1) * 4;
    const uint64_t by_l2 = std::max<uint64_t>(1, (uint64_t(512) << 10) / chunk_bytes);
    const uint64_t g0 = std::min<uint64_t>({uint64_t(nchunks), by_fill, by_l2});
    const uint64_t ngroups = (nchunks + g0 - 1) / g0;
    return uint32_t((nchunks + ngroups - 1) / ngroups);
}

// Short-launch eligibility (see gapblock_core.hpp) and its past-the-end prior value.
// T bounds |prefix sum| of one diagonal; with D = P0 - Pd in [-2T, 2T], big = INT32_MAX - 2T
// keeps D + big from wrapping, and 6T + prior_abs < INT32_MAX puts D + big above every real
// candidate D + prior <= 2T + prior_abs.
bool short_launch(uint32_t q_max, uint32_t r_max, uint32_t A, bool exact, int64_t pen_abs,
                  int64_t prior_abs, const PackedRefs& refs, int32_t* big) {
    const int64_t T = int64_t(std::max(q_max, r_max)) * pen_abs;
    if (exact || A > 31 || q_max > gb::kFastMaxL || r_max > gb::kShortMaxRef) return false;
    // the short kernels read the packed words through one buffer resource
    if (uint64_t(refs.nwords + 1) * refs.npad * 4 > gb::kBufferMaxBytes) return false;
    if (6 * T + prior_abs >= int64_t(INT32_MAX)) return false;
    *big = int32_t(int64_t(INT32_MAX) - 2 * T);
    return true;
}

GapblockPlan gapblock_plan(Arena& ar, const uint8_t* q, const uint64_t* q_off, uint64_t N,
                           uint32_t q_max, const uint8_t* r, const uint64_t* r_off, uint64_t K,
                           uint32_t r_max, const int32_t* pen_h, uint32_t A, int32_t go,
                           int32_t ge, const int32_t* prior_h, uint32_t W, bool pack,
                           hipStream_t s) {
    if (prior_h && W < std::max(q_max, r_max))
        throw HipError{ST_INVALID_ARGUMENT,
                       "seqtree_hip: prior_width " + std::to_string(W) +
                           " is below the longest sequence (" +
                           std::to_string(std::max(q_max, r_max)) + ")"};
    const DeviceTable pen_t = table_on_device(ar, pen_h, size_t(A) * A, 0, s);
    const int32_t* pen = pen_t.dev;
    const int64_t pen_abs = pen_t.abs_max;
    const size_t w1 = size_t(W) + 1;
    const int32_t* prior = nullptr;
    int64_t prior_abs = 0;
    if (prior_h) {
        // device prior cubes carry kPriorPad ints of zero padding (see gapblock_core.hpp)
        const DeviceTable pr = table_on_device(ar, prior_h, w1 * w1 * w1, gb::kPriorPad, s);
        prior = pr.dev;
        prior_abs = pr.abs_max;
    }

    GapblockPlan p{};
    GapblockArgs& a = p.a;
    a.q_codes = q;
    a.q_off = q_off;
    a.n_queries = N;
    a.n_refs = K;
    a.pen = pen;
    a.A = A;
    a.gap_open = go;
    a.gap_extend = ge;
    a.prior = prior;
    a.w1 = uint32_t(w1);
    // No int32 sum can wrap below this bound, so the one-pass minimum is bit-identical to
    // the reference's suffix/prefix evaluation.
    const int64_t Lmax = std::max(q_max, r_max);
    const int64_t worst = 2 * Lmax * pen_abs + prior_abs + int64_t(go) + Lmax * int64_t(ge);
    p.exact = worst >= int64_t(INT32_MAX);
    p.profile = size_t(q_max) * gb::kRowBytes <= 48 * 1024;  // one table row per position
    if (!pack) {
        a.group = 1;
        return p;
    }
    const PackedRefs geo = packed_geometry(K, r_max);
    p.short_path = short_launch(q_max, r_max, A, p.exact, pen_abs, prior_abs, geo, &a.big);
    TwinJob tw{};
    if (p.short_path) {
        a.prof_rows = q_max + gb::kFastMaxL;
        a.slice_rows = r_max + 1;
        p.twin = twin_launch(N, q_max, pen_abs, prior_abs, &a.big);
        if (p.twin) {
            // one scratch block: the twins (8-byte entries), then the length-sorted indices;
            // the packing launch builds the list
            const uint64_t cap = twin_capacity(N);
            p.twins = ar.get<uint2>(cap + (N + 1) / 2);
            tw = {q_off, N, reinterpret_cast<uint32_t*>(p.twins + cap), p.twins};
        }
    }
    a.refs = pack_refs(ar, r, r_off, K, r_max, s, 4, tw);
    // workgroups per chunk group: one per query (one per twin with twin rows)
    a.group = chunk_group(p.twin ? twin_capacity(N) : N, a.refs.nchunks, a.refs.nwords);
    return p;
}

void gapblock_run(const GapblockPlan& p, uint32_t q_max, hipStream_t s) {
    const std::string mode = p.a.hist         ? "gapblock_histogram"
                             : p.a.hits.count ? "gapblock_threshold"
                                              : "gapblock";
    if (p.twin) {
        LaunchTimer t(mode + "_twin", s);
        launch_gapblock_twin(p.a, p.twins, twin_capacity(p.a.n_queries), s);
        return;
    }
    LaunchTimer t(mode + (p.short_path ? "_short" : p.profile ? "_profile" : "_table") +
                      (p.exact ? "_exact" : ""),
                  s);
    launch_gapblock(p.a, p.profile, q_max, p.exact, p.short_path, s);
}

void gapblock_device(const uint8_t* q, const uint64_t* q_off, uint64_t N, uint32_t q_max,
                     const uint8_t* r, const uint64_t* r_off, uint64_t K, uint32_t r_max,
                     const int32_t* pen_h, uint32_t A, int32_t go, int32_t ge,
                     const int32_t* prior_h, uint32_t W, int32_t* out, hipStream_t s,
                     void* region = nullptr, size_t region_bytes = 0) {
    if (N == 0 || K == 0) return;
    Arena ar(s, region, region_bytes);
    GapblockPlan p = gapblock_plan(ar, q, q_off, N, q_max, r, r_off, K, r_max, pen_h, A, go, ge,
                                   prior_h, W, true, s);
    p.a.out = out;
    gapblock_run(p, q_max, s);
    launched("gapblock");
}

void gapblock_threshold_device(const uint8_t* q, const uint64_t* q_off, uint64_t N,
                               uint32_t q_max, const uint8_t* r, const uint64_t* r_off,
                               uint64_t K, uint32_t r_max, const int32_t* pen_h, uint32_t A,
                               int32_t go, int32_t ge, const int32_t* prior_h, uint32_t W,
                               const GapblockHits& hits, hipStream_t s) {
    launch_fill_u32(reinterpret_cast<uint32_t*>(hits.count), 2, 0u, s);
    if (N == 0 || K == 0) return;
    Arena ar(s);
    GapblockPlan p = gapblock_plan(ar, q, q_off, N, q_max, r, r_off, K, r_max, pen_h, A, go, ge,
                                   prior_h, W, true, s);
    p.a.hits = hits;
    gapblock_run(p, q_max, s);
    launched("gapblock_threshold");
}

void gapblock_pairs_device(const uint8_t* q, const uint64_t* q_off, uint64_t N, uint32_t q_max,
                           const uint8_t* r, const uint64_t* r_off, uint64_t K, uint32_t r_max,
                           const int32_t* pen_h, uint32_t A, int32_t go, int32_t ge,
                           const int32_t* prior_h, uint32_t W, const uint32_t* pq,
                           const uint32_t* pr, uint64_t n_pairs, int32_t* out, hipStream_t s) {
    if (n_pairs == 0) return;
    Arena ar(s);
    GapblockPlan p = gapblock_plan(ar, q, q_off, N, q_max, r, r_off, K, r_max, pen_h, A, go, ge,
                                   prior_h, W, false, s);
    p.a.out = out;
    {
        LaunchTimer t(std::string("gapblock_pairs") + (p.exact ? "_exact" : ""), s);
        launch_gapblock_pairs(p.a, r, r_off, pq, pr, n_pairs, p.exact, s);
    }
    launched("gapblock_pairs");
}

// ------------------------------------------------------------------------------------------
// ragged multi-job gap-block batch (st_gapblock_batch_*)
// ------------------------------------------------------------------------------------------
struct DeviceBuffer {
    void* p = nullptr;
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
};

struct BatchPlanImpl {
    std::vector<st_gapblock_job> jobs;        // caller's descriptors (device sequences/outputs)
    std::vector<std::unique_ptr<DeviceBuffer>> owned;
    std::vector<BatchJob> dev_jobs;           // in unit order
    std::vector<uint32_t> order;              // dev_jobs[t] is jobs[order[t]]
    std::vector<PackedRefs> refs;             // per caller job
    struct PackArgs {
        const uint8_t* bytes;
        const uint64_t* off;
        uint64_t n;
        uint32_t nwords, npad;
        uint32_t *words, *len, *orig;
        uint32_t scale, max_len;
        TwinJob tw;  // the job's twin list, built by the same launch (twin jobs)
    };
    std::vector<PackArgs> packs;
    BatchJob* d_jobs = nullptr;
    unsigned long long* d_next = nullptr;
    uint64_t total_units = 0;
    size_t lds_bytes = 0;        // stage + the largest job's tables
    uint32_t stage_ints = kChunk;  // 2 kChunk when some job has twin rows
    bool any_twin = false;
    bool all_twin = false;  // every job has twin rows: the twin-only kernel
    BatchTiles tiles{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int cus = 256;

    ~BatchPlanImpl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    template <class T>
    T* alloc(size_t count) {
        auto b = std::make_unique<DeviceBuffer>();
        check(hipMalloc(&b->p, std::max<size_t>(count * sizeof(T), 256)), "hipMalloc");
        T* out = static_cast<T*>(b->p);
        owned.push_back(std::move(b));
        return out;
    }
    // Orders units heaviest job first and uploads the descriptors for this tile shape.
    void layout(uint32_t qt, hipStream_t s) {
        uint64_t u = 0;
        for (auto& j : dev_jobs) {
            j.unit_begin = u;
            const uint64_t rows = j.twin ? twin_capacity(j.n_queries) : j.n_queries;
            j.qblocks = uint32_t((rows + qt - 1) / qt);
            const uint32_t ngroups = (j.refs.nchunks + j.group - 1) / j.group;
            u += uint64_t(j.qblocks) * ngroups;
        }
        total_units = u;
        h2d(d_jobs, dev_jobs.data(), dev_jobs.size(), s);
    }
    // The whole batch as stream work: pack every job's references, reset the unit counter
    // (a fill kernel: no memset nodes in a captured graph), run the persistent kernel.
    void enqueue(hipStream_t s) {
        for (const auto& p : packs)
            launch_pack_refs(p.bytes, p.off, p.n, p.nwords, p.npad, p.words, p.len, p.orig,
                             p.scale, p.max_len, p.tw, s);
        launch_fill_u32(reinterpret_cast<uint32_t*>(d_next), 2, 0u, s);
        launch_gapblock_batch(d_jobs, uint32_t(dev_jobs.size()), total_units, tiles, stage_ints,
                              lds_bytes, all_twin, d_next, s);
    }
};

// Tile shapes the autotuner tries (queries per unit, threads per workgroup, workgroups/CU).
constexpr uint32_t kTuneQpu[] = {1, 4, 16};
constexpr uint32_t kTuneThreads[] = {256, 512};
constexpr uint32_t kTuneWgPerCu[] = {4, 8};

// Process-wide memo of tuned shapes, keyed by a workload signature.
std::mutex g_tune_mu;
std::map<std::string, BatchTiles> g_tuned;

std::string batch_signature(const std::vector<st_gapblock_job>& jobs) {
    // job count, total queries/refs (log2 buckets) and the widest lengths
    uint64_t nq = 0, nr = 0;
    uint32_t lq = 0, lr = 0;
    for (const auto& j : jobs) {
        nq += j.n_queries;
        nr += j.n_refs;
        lq = std::max(lq, j.q_max_len);
        lr = std::max(lr, j.r_max_len);
    }
    auto lg = [](uint64_t v) { int b = 0; while (v >>= 1) ++b; return b; };
    return std::to_string(jobs.size()) + ":" + std::to_string(lg(nq)) + ":" +
           std::to_string(lg(nr)) + ":" + std::to_string(lq) + ":" + std::to_string(lr);
}

// ------------------------------------------------------------------------------------------
// Gotoh
// ------------------------------------------------------------------------------------------
// Register-kernel row classes: a query of length m runs in the class R = roundup(m, 4).
constexpr int kGotohClasses = 16;
// Long queries run one wave per pair (gotoh_wave_kernel) up to this many pairs.
constexpr uint64_t kWavePairsMax = uint64_t(1) << 18;
constexpr int gotoh_rows(int c) { return 4 * (c + 1); }

void gotoh_device(const uint8_t* q, const uint64_t* q_off, uint64_t N, uint32_t q_max,
                  const uint8_t* r, const uint64_t* r_off, uint64_t K, uint32_t r_max,
                  const int32_t* sim_h, uint32_t A, int local, int dist, int32_t go, int32_t ge,
                  int32_t* out, hipStream_t s, bool* overflow_out) {
    *overflow_out = false;
    if (N == 0 || K == 0) return;
    Arena ar(s);
    std::vector<int32_t> simT_h(size_t(A) * A);
    int64_t mx = std::max<int64_t>(go, ge);
    for (uint32_t x = 0; x < A; ++x)
        for (uint32_t y = 0; y < A; ++y) {
            simT_h[size_t(y) * A + x] = sim_h[size_t(x) * A + y];
            mx = std::max(mx, iabs64(sim_h[size_t(x) * A + y]));
        }
    int32_t* sim = ar.get<int32_t>(size_t(A) * A);
    int32_t* simT = ar.get<int32_t>(size_t(A) * A);
    h2d(sim, sim_h, size_t(A) * A, s);
    h2d(simT, simT_h.data(), size_t(A) * A, s);
    int32_t* flag = ar.get<int32_t>(1);
    launch_fill_u32(reinterpret_cast<uint32_t*>(flag), 1, 0, s);

    GotohArgs a{};
    a.q_codes = q;
    a.q_off = q_off;
    a.r_codes = r;
    a.r_off = r_off;
    a.n_queries = N;
    a.n_refs = K;
    a.sim = sim;
    a.simT = simT;
    a.A = A;
    a.local = local;
    a.dist = dist;
    a.gap_open = go;
    a.gap_extend = ge;
    a.out = out;
    a.overflow = flag;

    // Per-thread int64 scratch for the generic kernels: 6 rows over the shorter sequence.
    const uint64_t shorter = std::min(q_max, r_max);
    const uint64_t self_len = std::max(q_max, r_max);
    auto slots_for = [](uint64_t stride) {
        const uint64_t per = 6 * stride * sizeof(int64_t);
        uint64_t slots = (uint64_t(1) << 30) / std::max<uint64_t>(per, 1);
        slots = std::min<uint64_t>(slots, 65536);
        slots = std::max<uint64_t>(slots / 256 * 256, 256);
        return uint32_t(slots);
    };

    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // the intra-pair wavefront kernel: int32 cells in the H form (see launch_gotoh_wave)
    auto wave_fits = [&](uint32_t len_q, uint32_t len_r) {
        return go >= ge && (int64_t(len_q) + len_r + 2) * mx < (int64_t(1) << 28);
    };
    auto run_wave = [&](const GotohWaveJob& job, uint32_t len_q, uint32_t len_r,
                        const char* label) {
        const uint32_t grid = uint32_t(std::min<uint64_t>((job.n_pairs + 3) / 4,
                                                          uint64_t(cus) * 4));
        int32_t* bnd = nullptr;
        if (const uint64_t words = gotoh_wave_bnd_words(grid, len_q, len_r))
            bnd = ar.get<int32_t>(words);
        LaunchTimer t(std::string(label) + (local ? "_local" : "_global"), s);
        launch_gotoh_wave(a, job, len_q, len_r, grid, bnd, s);
    };

    if (dist) {
        int64_t* qs = ar.get<int64_t>(N);
        int64_t* rs = ar.get<int64_t>(K);
        if (gotoh_self_fits_lds(uint32_t(self_len),
                                std::max({mx, iabs64(go), iabs64(ge)}))) {
            LaunchTimer t("gotoh_self", s);
            const SelfSets sets{{q, r}, {q_off, r_off}, {N, K}, {qs, rs}, 0};
            launch_gotoh_self_lds(a, sets, uint32_t(self_len), s);
        } else if (wave_fits(uint32_t(self_len), uint32_t(self_len))) {
            // long strings: one wave per string
            const uint32_t L = uint32_t(self_len);
            run_wave(GotohWaveJob{nullptr, N + K, L, L, {q, r}, {q_off, r_off}, {qs, rs}, N}, L,
                     L, "gotoh_wave_self");
        } else {
            LaunchTimer t("gotoh_self", s);
            const uint64_t stride = self_len + 1;
            const uint32_t slots = slots_for(stride);
            int64_t* scratch = ar.get<int64_t>(size_t(slots) * 6 * stride);
            launch_gotoh_self(a, q, q_off, N, qs, scratch, slots, stride, s);
            launch_gotoh_self(a, r, r_off, K, rs, scratch, slots, stride, s);
        }
        launched("gotoh_self");
        a.q_self = qs;
        a.r_self = rs;
    }

    // Range proof for the register kernels: with R <= 64 stripe rows (dummies included), or
    // the query swept in stripes of 64 rows up to its padded length q_pad, every reachable
    // cell is a partial-alignment score bounded by B = (max(64, q_pad) + r_max) * mx; B, the
    // sentinel (-2^30 in int32, -2^14 in int16) and one gap step below either must stay
    // inside the cell type. A distance adds at most 4 * (q_max + r_max) * mx, formed in
    // int64 and stored as int32.
    const int64_t q_pad = std::max<int64_t>(
        64, (int64_t(q_max) + kStripeRows - 1) / kStripeRows * kStripeRows);
    const int64_t B = (q_pad + int64_t(r_max)) * mx;
    const int64_t G = std::max<int64_t>(go, ge);
    bool fast = B + 2 * G < (int64_t(1) << 30);
    // the packed cell uses the H form only, which needs gap_open >= gap_extend
    const bool packed16 = B + 2 * G < (int64_t(1) << 14) && go >= ge;
    if (dist) fast = fast && 4 * (int64_t(q_max) + int64_t(r_max)) * mx < int64_t(INT32_MAX);

    ClassBounds cb{};
    uint32_t* lists = nullptr;
    std::vector<uint32_t> counts;
    if (fast) {
        cb.count = kGotohClasses + 1;
        for (int c = 0; c < kGotohClasses; ++c) cb.bound[c] = uint32_t(gotoh_rows(c));
        cb.bound[kGotohClasses] = 0xFFFFFFFFu;
    } else {
        cb.count = 1;
        cb.bound[0] = 0xFFFFFFFFu;
    }
    counts = classify(ar, q_off, N, cb, &lists, s);

    bool generic_ran = false;
    if (fast) {
        a.refs = pack_refs(ar, r, r_off, K, r_max, s);
        for (int c = 0; c < kGotohClasses; ++c) {
            if (counts[c] == 0) continue;
            {
                LaunchTimer t(std::string(packed16 ? "gotoh_i16x2_r" : "gotoh_i32_r") +
                                  std::to_string(gotoh_rows(c)) + (local ? "_local" : "_global"),
                              s);
                launch_gotoh_fast(a, gotoh_rows(c), packed16, lists + size_t(c) * N, counts[c], s);
            }
            launched("gotoh_fast");
        }
    }
    const int gen_class = cb.count - 1;
    // Few long queries: a wave per pair rather than a lane per pair (the multi-stripe kernel
    // below gives a lane to each pair and needs about 2^18 pairs to fill the chip).
    const bool wave = fast && counts[gen_class] > 0 && wave_fits(q_max, r_max) &&
                      uint64_t(counts[gen_class]) * K <= kWavePairsMax;
    if (wave) {
        run_wave(GotohWaveJob{lists + size_t(gen_class) * N, uint64_t(counts[gen_class]) * K,
                              q_max, r_max, {}, {}, {}, 0},
                 q_max, r_max, "gotoh_wave");
        launched("gotoh_wave");
    } else if (fast && counts[gen_class] > 0) {
        // queries past one register stripe: the multi-stripe register kernel
        const uint32_t n_long = counts[gen_class];
        const bool hform = packed16 || go >= ge;
        const uint32_t W = std::max<uint32_t>(r_max, 1);
        const uint64_t units = gotoh_stripes_units(n_long, packed16, a.refs.nchunks);
        const uint64_t per_wg = gotoh_stripes_words(1, W, hform) * sizeof(uint32_t);
        const uint64_t by_mem = std::max<uint64_t>(1, (uint64_t(1) << 30) / per_wg);
        const uint32_t grid =
            uint32_t(std::min<uint64_t>({units, uint64_t(cus) * 8, by_mem}));
        uint32_t* bnd = ar.get<uint32_t>(gotoh_stripes_words(grid, W, hform));
        {
            LaunchTimer t(std::string(packed16 ? "gotoh_i16x2_stripes" : "gotoh_i32_stripes") +
                              (local ? "_local" : "_global"),
                          s);
            launch_gotoh_stripes(a, packed16, lists + size_t(gen_class) * N, n_long, grid, bnd,
                                 W, q_max, s);
        }
        launched("gotoh_stripes");
    } else if (counts[gen_class] > 0) {
        const uint64_t stride = shorter + 1;
        const uint32_t slots = slots_for(stride);
        int64_t* scratch = ar.get<int64_t>(size_t(slots) * 6 * stride);
        {
            LaunchTimer t("gotoh_generic", s);
            launch_gotoh_generic(a, lists + size_t(gen_class) * N, counts[gen_class], scratch,
                                 slots, stride, s);
        }
        launched("gotoh_generic");
        generic_ran = true;
    }
    if (generic_ran) {
        int32_t h = 0;
        d2h(&h, flag, 1, s);
        sync(s);
        *overflow_out = h != 0;
    }
}

// ------------------------------------------------------------------------------------------
// Hamming / Levenshtein
// ------------------------------------------------------------------------------------------
void hamming_device(const uint8_t* a_b, const uint64_t* a_off, uint64_t N, uint32_t a_max,
                    const uint8_t* b_b, const uint64_t* b_off, uint64_t K, uint32_t b_max,
                    int32_t* out, hipStream_t s) {
    if (N == 0 || K == 0) return;
    Arena ar(s);
    const uint32_t L = std::max(a_max, b_max);  // all lengths are equal by now
    const uint32_t nw = words_for(L);
    DistanceArgs d{};
    d.a_bytes = a_b;
    d.a_off = a_off;
    d.b_bytes = b_b;
    d.b_off = b_off;
    d.n_a = N;
    d.n_b = K;
    d.out = out;
    d.refs = pack_refs(ar, b_b, b_off, K, L, s);
    uint32_t* aw = ar.get<uint32_t>(size_t(N) * std::max<uint32_t>(nw, 1));
    launch_pack_rows(a_b, a_off, N, nw, aw, s);
    {
        LaunchTimer t("hamming", s);
        launch_hamming(d, aw, nw, s);
    }
    launched("hamming");
}

void levenshtein_device(const uint8_t* a_b, const uint64_t* a_off, uint64_t N, uint32_t a_max,
                        const uint8_t* b_b, const uint64_t* b_off, uint64_t K, uint32_t b_max,
                        int32_t* out, hipStream_t s) {
    if (N == 0 || K == 0) return;
    Arena ar(s);
    DistanceArgs d{};
    d.a_bytes = a_b;
    d.a_off = a_off;
    d.b_bytes = b_b;
    d.b_off = b_off;
    d.n_a = N;
    d.n_b = K;
    d.out = out;
    ClassBounds cb{};
    cb.count = 5;
    cb.bound[0] = 64;
    cb.bound[1] = 128;
    cb.bound[2] = 192;
    cb.bound[3] = 256;
    cb.bound[4] = 0xFFFFFFFFu;
    uint32_t* lists = nullptr;
    const std::vector<uint32_t> counts = classify(ar, a_off, N, cb, &lists, s);
    if (counts[0] + counts[1] + counts[2] + counts[3] > 0) {
        d.refs = pack_refs(ar, b_b, b_off, K, b_max, s);
        for (int c = 0; c < 4; ++c) {
            if (counts[c] == 0) continue;
            {
                LaunchTimer t("levenshtein_myers_w" + std::to_string(c + 1), s);
                launch_levenshtein_myers(d, c + 1, lists + size_t(c) * N, counts[c], s);
            }
            launched("levenshtein_myers");
        }
    }
    if (counts[4] > 0) {
        const uint64_t stride = uint64_t(std::min(a_max, b_max)) + 1;
        uint64_t slots = (uint64_t(1) << 30) / (2 * stride * sizeof(int32_t));
        slots = std::max<uint64_t>(std::min<uint64_t>(slots, 65536) / 256 * 256, 256);
        int32_t* scratch = ar.get<int32_t>(size_t(slots) * 2 * stride);
        LaunchTimer t("levenshtein_generic", s);
        launch_levenshtein_generic(d, lists + size_t(4) * N, counts[4], scratch, uint32_t(slots),
                                   stride, s);
        launched("levenshtein_generic");
    }
}

// ------------------------------------------------------------------------------------------
// host-memory plumbing
// ------------------------------------------------------------------------------------------
struct HostSet {
    const uint8_t* bytes;
    const uint64_t* off;
    uint64_t n;
    uint32_t max_len;
    uint64_t total;
};

HostSet host_set(const uint8_t* bytes, const uint64_t* off, uint64_t n) {
    HostSet h{bytes, off, n, 0, n ? off[n] : 0};
    for (uint64_t i = 0; i < n; ++i) h.max_len = std::max<uint32_t>(h.max_len, uint32_t(off[i + 1] - off[i]));
    return h;
}

struct DeviceSet {
    const uint8_t* bytes;
    const uint64_t* off;
};

DeviceSet upload(Arena& ar, const HostSet& h, HostCall& call) {
    uint8_t* b = ar.get<uint8_t>(std::max<uint64_t>(h.total, 1));
    uint64_t* o = ar.get<uint64_t>(h.n + 1);
    call.put(b, h.bytes, h.total);
    call.put(o, h.off, h.n + 1);
    return {b, o};
}

// Runs `body(q_dev_subset, row0, rows, out_dev)` over blocks of query rows and copies each
// block of the matrix to `out` while the next block is scored: the host enqueues block b, then
// copies block b - 1 out (pinned staging on the copy stream, CopyPool into `out`). Blocks hold
// about kRowBlockBytes of output, so a small matrix is one block; the device keeps at most
// 8 GiB of output at a time.
constexpr uint64_t kRowBlockBytes = uint64_t(128) << 20;

template <class Body>
void run_rows(const HostSet& q, uint64_t K, int32_t* out, HostCall& call, Arena& ar,
              const DeviceSet& qd, Body&& body) {
    const uint64_t row_bytes = std::max<uint64_t>(K * 4, 1);
    const uint64_t cap = std::min<uint64_t>(std::max<uint64_t>(1, (uint64_t(8) << 30) / row_bytes),
                                            q.n);
    const uint64_t n_blocks = std::max<uint64_t>(1, (cap * row_bytes + kRowBlockBytes - 1) /
                                                        kRowBlockBytes);
    const uint64_t block = (cap + n_blocks - 1) / n_blocks;
    int32_t* dout = ar.get<int32_t>(size_t(cap) * K);
    for (uint64_t r0 = 0; r0 < q.n; r0 += cap) {
        const uint64_t rows = std::min(cap, q.n - r0);
        hipEvent_t prev = nullptr;
        uint64_t prev_b0 = 0, prev_nb = 0;
        for (uint64_t b0 = 0; b0 < rows; b0 += block) {
            const uint64_t nb = std::min(block, rows - b0);
            body(qd.off + r0 + b0, r0 + b0, nb, dout + b0 * K);
            hipEvent_t done = call.mark();
            if (prev)
                call.download_after(prev, out + (r0 + prev_b0) * K, dout + prev_b0 * K,
                                    size_t(prev_nb) * K * 4);
            prev = done;
            prev_b0 = b0;
            prev_nb = nb;
        }
        call.download_after(prev, out + (r0 + prev_b0) * K, dout + prev_b0 * K,
                            size_t(prev_nb) * K * 4);
        call.finish();  // the next rows reuse dout
    }
}

template <class F>
int guarded(st_status* st, F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        set_status(st, e.code, "%s", e.msg.c_str());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_status(st, ST_OUT_OF_MEMORY, "seqtree_hip: host allocation failed");
        return ST_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_status(st, ST_DEVICE_ERROR, "seqtree_hip: %s", e.what());
        return ST_DEVICE_ERROR;
    }
}

const char* overflow_message(bool dist) {
    return dist ? "seqtree: the alignment distance does not fit in 32 bits -- the gap costs are "
                  "absurdly large for these sequences"
                : "seqtree: the alignment score does not fit in 32 bits -- the gap costs are "
                  "absurdly large for these sequences";
}

int bad_alphabet(st_status* st, uint32_t A) {
    set_status(st, ST_INVALID_ARGUMENT, "seqtree_hip: alphabet size %u is outside 1..32", A);
    return ST_INVALID_ARGUMENT;
}

// First length-mismatched (i, k) in row-major order, or false.
bool first_mismatch(const HostSet& a, const HostSet& b, uint64_t* i, uint64_t* k) {
    if (a.n == 0 || b.n == 0) return false;
    const uint64_t la0 = a.off[1] - a.off[0];
    for (uint64_t t = 0; t < b.n; ++t)
        if (b.off[t + 1] - b.off[t] != la0) {
            *i = 0;
            *k = t;
            return true;
        }
    for (uint64_t t = 0; t < a.n; ++t)
        if (a.off[t + 1] - a.off[t] != la0) {
            *i = t;
            *k = 0;
            return true;
        }
    return false;
}

int hamming_mismatch(st_status* st, uint64_t la, uint64_t lb) {
    set_status(st, ST_INVALID_ARGUMENT,
               "seqtree: hamming distance is only defined for equal-length sequences (got %llu "
               "and %llu)",
               (unsigned long long)la, (unsigned long long)lb);
    return ST_INVALID_ARGUMENT;
}


// ------------------------------------------------------------------------------------------
// learned prefilter (st_prefilter_*)
// ------------------------------------------------------------------------------------------
// Activation buffers: R rows (a multiple of the 128-row GEMM tile) plus kGuardRows zero rows
// on either side, so the five-row convolution window of every row stays in bounds.
constexpr uint32_t kGuardRows = 2;

struct PfActs {
    __bf16* base = nullptr;  // allocation start
    __bf16* rows = nullptr;  // row 0
};

PfActs pf_acts(Arena& ar, uint64_t R, uint32_t C, hipStream_t s,
               uint32_t guard = kGuardRows) {
    PfActs a;
    const uint64_t total = (R + 2 * uint64_t(guard)) * C;
    a.base = ar.get<__bf16>(total);
    a.rows = a.base + uint64_t(guard) * C;
    // zero the guard rows (bf16 zero is all-zero bits; two elements per 32-bit word)
    launch_fill_u32(reinterpret_cast<uint32_t*>(a.base), uint64_t(guard) * C / 2, 0u, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(a.rows + R * C), uint64_t(guard) * C / 2, 0u, s);
    return a;
}

// Rows of a packed set: every sequence plus its two zero rows, rounded up to the GEMM tile.
struct PfRows {
    uint64_t used;
    uint64_t R;
};

PfRows pf_rows(uint64_t off_first, uint64_t off_last, uint32_t n) {
    const uint64_t used = (off_last - off_first) + 2ull * n;
    return {used, (used + 127) / 128 * 128};
}

uint64_t read_offset(const uint64_t* off, uint32_t i, hipStream_t s) {
    uint64_t h = 0;
    d2h(&h, off + i, 1, s);
    sync(s);
    return h;
}

// Encoder forward over one packed block of sequences: X0 (and `valid`) for every row, then
// the three convolutions; H3 is left in h3.
void pf_forward(const __bf16* B, const float* P, const uint8_t* codes, const uint64_t* off,
                uint32_t n, const PfRows& rows, const PfActs& x0, const PfActs& h1,
                const PfActs& h2, const PfActs& h3, uint8_t* valid, hipStream_t s) {
    LaunchTimer t("prefilter_encode", s);
    const uint64_t R = rows.R;
    launch_fill_u32(reinterpret_cast<uint32_t*>(valid), R / 4, 0u, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(x0.rows + rows.used * pf::kEmbDim),
                    (R - rows.used) * pf::kEmbDim / 2, 0u, s);
    pf::launch_embed(codes, off, n, B + pf::kOffEmbB, x0.rows, valid, nullptr, s);
    pf::launch_conv(1, 0, x0.rows, B + pf::kOffWt1, P + pf::kOffB1, nullptr, valid, h1.rows, R, s);
    pf::launch_conv(2, 0, h1.rows, B + pf::kOffWt2, P + pf::kOffB2, nullptr, valid, h2.rows, R, s);
    pf::launch_conv(3, 0, h2.rows, B + pf::kOffWt3, P + pf::kOffB3, nullptr, valid, h3.rows, R, s);
    launched("prefilter forward");
}

// Scratch of the fused encoder for blocks of up to R rows.
struct PfEncodeScratch {
    uint32_t* tile_first;
    float* part;
    uint32_t* head_seq;
};

PfEncodeScratch pf_encode_scratch(Arena& ar, uint64_t R) {
    const uint64_t tiles = R / 128;
    return {ar.get<uint32_t>(tiles), ar.get<float>(tiles * 2 * pf::kChannels),
            ar.get<uint32_t>(tiles)};
}

// Guard rows of the fused encoder's inputs: its tile windows read X0 rows -6 .. R + 5 and
// valid bytes -4 .. R + 43, all zero outside the block's rows.
constexpr uint32_t kEncX0Guard = 8;     // X0 rows before row 0 and past R
constexpr uint32_t kEncValidHead = 16;  // valid bytes before row 0
constexpr uint32_t kEncValidTail = 64;  // valid bytes past R

// Inference forward and mean pool over one packed block: X0, `valid` and the tiles' first
// sequences, then the fused kernel (H1 .. H3 never leave LDS) writing the BF16 embeddings eb
// and their squared norms. x0 has kEncX0Guard guard rows on each side; valid points
// kEncValidHead bytes into an allocation of R + kEncValidHead + kEncValidTail bytes.
void pf_encode_forward(const __bf16* B, const float* P, const uint8_t* codes, const uint64_t* off,
                       uint32_t n, const PfRows& rows, const PfActs& x0, uint8_t* valid,
                       const PfEncodeScratch& sc, __bf16* eb, float* norm, hipStream_t s) {
    LaunchTimer t("prefilter_encode", s);
    const uint64_t R = rows.R;
    launch_fill_u32(reinterpret_cast<uint32_t*>(valid - kEncValidHead),
                    (R + kEncValidHead + kEncValidTail) / 4, 0u, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(x0.rows + rows.used * pf::kEmbDim),
                    (R + kEncX0Guard - rows.used) * pf::kEmbDim / 2, 0u, s);
    launch_fill_u32(sc.tile_first, R / 128, n, s);
    pf::launch_embed(codes, off, n, B + pf::kOffEmbB, x0.rows, valid, sc.tile_first, s);
    pf::launch_encode(x0.rows, valid, B, P, R, off, n, sc.tile_first, eb, norm, sc.part,
                      sc.head_seq, s);
    launched("prefilter encode forward");
}

// ------------------------------------------------------------------------------------------
// Gotoh alignment with traceback of a pair list
// ------------------------------------------------------------------------------------------
// Runs the traceback kernels on device-resident sets: int32 cells when every value stays far
// inside int32 ((q_max + r_max + 2) * max(|sim|, go, ge) < 2^28, so the -2^29 sentinel and a
// few gap steps below it keep every comparison of the int64 DP), int64 cells otherwise. The
// direction bytes (and, past the LDS limit or for int64, the DP rows) take per-workgroup
// scratch; the grid is sized so that scratch stays under 1 GiB. Lists of 4,096 pairs or more
// are processed in shape order (a counting sort by lengths clamped to 31), so a wave's lanes
// run the same number of cells; every output still goes to the pair's own index.
//
// The penalty form (Index::align, src/index.cpp:105-179) runs the same kernels in global mode
// on the negated penalty table (sim_h holds -penalty) with minus the reference's int32 kInf as
// the int64 cells' -inf, and writes the penalty as int32 to score32 instead of score.
//
// Each pair's lengths are checked against q_max / r_max on the device: a longer pair is
// skipped and the call fails with ST_INVALID_ARGUMENT after one stream synchronisation (the
// direction scratch is sized by the bounds).
void traceback_device(const uint8_t* q, const uint64_t* q_off, uint32_t q_max, const uint8_t* r,
                      const uint64_t* r_off, uint32_t r_max, const uint32_t* pq,
                      const uint32_t* pr, uint64_t n_pairs, const int32_t* sim_h, uint32_t A,
                      int32_t local, int32_t go, int32_t ge, uint8_t* ops, uint64_t ops_stride,
                      int64_t* score, uint32_t* n_ops, uint32_t* q_begin, uint32_t* r_begin,
                      hipStream_t s, int32_t* score32 = nullptr) {
    if (go < 0 || ge < 0)
        throw HipError{ST_INVALID_ARGUMENT, "seqtree: gap_open and gap_extend must be >= 0"};
    if (ops_stride < uint64_t(q_max) + r_max)
        throw HipError{ST_INVALID_ARGUMENT,
                       "seqtree_hip: ops_stride is below q_max_len + r_max_len"};
    if (n_pairs == 0) return;
    const bool penalty = score32 != nullptr;
    Arena ar(s);
    uint32_t* too_long = ar.get<uint32_t>(1);
    launch_fill_u32(too_long, 1, 0u, s);
    const DeviceTable sim = table_on_device(ar, sim_h, size_t(A) * A, 0, s);
    const int64_t mx = std::max<int64_t>({sim.abs_max, int64_t(go), int64_t(ge)});
    const bool wide = (int64_t(q_max) + r_max + 2) * mx >= (int64_t(1) << 28);
    // registers for the rows of short pairs, LDS rows up to 64 reference residues, else scratch
    const bool reg = !wide && r_max <= kTbRegCols && q_max <= kTbRegMaxQ;
    // four bits per cell reproduce every back-pointer when gap_open >= gap_extend
    const bool nib = reg && go >= ge;
    const bool lds_rows = !wide && r_max <= kTracebackLdsMaxLen;
    const uint64_t W = uint64_t(r_max) + 1;
    const uint64_t dir_bytes = nib   ? std::max<uint64_t>(q_max, 1) * kTbNibDwords * 64 * 4
                               : reg ? (uint64_t(q_max) + 1) * kTbRegGroups * 64 * 4
                                     : (uint64_t(q_max) + 1) * W * 64;  // per workgroup
    const uint64_t row_bytes = lds_rows ? 0 : 3 * W * 64 * (wide ? 8 : 4);
    const uint64_t need = (n_pairs + 63) / 64;
    const uint64_t fit = std::max<uint64_t>(1, (uint64_t(1) << 30) / (dir_bytes + row_bytes));
    const uint32_t wgs = uint32_t(std::min<uint64_t>({need, fit, 65535}));
    uint8_t* dirs = ar.get<uint8_t>(dir_bytes * wgs);
    void* rows = row_bytes ? static_cast<void*>(ar.get<uint8_t>(row_bytes * wgs)) : nullptr;
    TracebackArgs a{};
    a.q_codes = q;
    a.q_off = q_off;
    a.r_codes = r;
    a.r_off = r_off;
    a.pair_q = pq;
    a.pair_r = pr;
    a.n_pairs = n_pairs;
    a.sim = sim.dev;
    a.A = A;
    a.local = local;
    a.gap_open = go;
    a.gap_extend = ge;
    a.q_max = q_max;
    a.r_max = r_max;
    a.ops = ops;
    a.ops_stride = ops_stride;
    a.score = score;
    a.n_ops = n_ops;
    a.q_begin = q_begin;
    a.r_begin = r_begin;
    a.neg_inf = penalty ? -int64_t(INT32_MAX / 4) : INT64_MIN / 4;
    a.score32 = score32;
    a.too_long = too_long;
    if (n_pairs >= 4096 && n_pairs < (uint64_t(1) << 32)) {
        // lanes of one wave take pairs of one shape (the DP and the walk run per lane)
        const uint64_t table_n =
            (uint64_t(traceback_order_slices(n_pairs)) + 1) * kTracebackShapeBins;
        uint32_t* table = ar.get<uint32_t>(table_n + 2 * n_pairs);
        uint32_t* order = table + table_n + n_pairs;
        {
            LaunchTimer t("traceback_order", s);
            launch_traceback_order(a, table, table + table_n, order, s);
        }
        launched("traceback_order");
        a.order = order;
    }
    {
        LaunchTimer t(std::string("traceback_") +
                          (nib ? "nib" : reg ? "reg" : wide ? "i64" : "i32") +
                          (penalty ? "_penalty" : local ? "_local" : "_global"),
                      s);
        if (nib) launch_traceback_nib(a, wgs, reinterpret_cast<uint32_t*>(dirs), s);
        else if (reg) launch_traceback_reg(a, wgs, reinterpret_cast<uint32_t*>(dirs), s);
        else launch_traceback(a, wide, lds_rows, wgs, rows, dirs, s);
    }
    launched("traceback");
    uint32_t flag = 0;
    d2h(&flag, too_long, 1, s);
    sync(s);
    if (flag)
        throw HipError{ST_INVALID_ARGUMENT,
                       "seqtree_hip: a pair is longer than q_max_len / r_max_len"};
}

// The negated penalty table of Index::align: -penalty[a][b], or -(a != b) for unit cost.
std::vector<int32_t> negated_penalties(const int32_t* pen, uint32_t A) {
    std::vector<int32_t> t(size_t(A) * A);
    for (uint32_t x = 0; x < A; ++x)
        for (uint32_t y = 0; y < A; ++y)
            t[size_t(x) * A + y] = pen ? int32_t(0u - uint32_t(pen[size_t(x) * A + y]))
                                       : -int32_t(x != y);
    return t;
}

// ------------------------------------------------------------------------------------------
// Per-query histograms of exact gap-block scores (control-calibrated counts)
// ------------------------------------------------------------------------------------------
// The dense kernels score blocks of query rows into a scratch block (at most 512 MB), and a
// histogram pass folds each row into out[row][H + 2] (uint64; bin 0 below zero, 1 + s for
// 0 <= s < H, H + 1 at H or above): the count of references within any cutoff below H and
// every order statistic of the scores follow from it.
void gapblock_histogram_device(const uint8_t* q, const uint64_t* q_off, uint64_t N,
                               uint32_t q_max, const uint8_t* r, const uint64_t* r_off,
                               uint64_t K, uint32_t r_max, const int32_t* pen_h, uint32_t A,
                               int32_t go, int32_t ge, const int32_t* prior_h, uint32_t W,
                               uint32_t H, uint64_t* out, hipStream_t s) {
    if (H == 0 || uint64_t(H) + 2 > kMaxHistBins)
        throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: histogram bins must be 1.." +
                                                std::to_string(kMaxHistBins - 2)};
    if (N == 0) return;
    launch_fill_u32(reinterpret_cast<uint32_t*>(out), 2 * N * (uint64_t(H) + 2), 0u, s);
    if (K == 0) return;
    Arena ar(s);
    if (H + 2 <= kHistFusedBins) {
        // fused: the scoring kernels add each score to an LDS histogram of their query rows,
        // no score matrix is written
        GapblockPlan p = gapblock_plan(ar, q, q_off, N, q_max, r, r_off, K, r_max, pen_h, A, go,
                                       ge, prior_h, W, true, s);
        p.a.hist = reinterpret_cast<unsigned long long*>(out);
        p.a.hist_bins = H;
        gapblock_run(p, q_max, s);
        launched("gapblock_histogram");
        return;
    }
    // wider histograms: dense row blocks, then one reduction pass per block
    const uint64_t blk = std::min<uint64_t>({N, std::max<uint64_t>(1, (uint64_t(512) << 20) / (4 * K)),
                                             65535});
    int32_t* dense = ar.get<int32_t>(blk * K);
    for (uint64_t r0 = 0; r0 < N; r0 += blk) {
        const uint64_t rows = std::min(blk, N - r0);
        gapblock_device(q, q_off + r0, rows, q_max, r, r_off, K, r_max, pen_h, A, go, ge, prior_h,
                        W, dense, s);
        LaunchTimer t("score_histogram", s);
        launch_score_histogram(dense, rows, K, H,
                               reinterpret_cast<unsigned long long*>(out + r0 * (uint64_t(H) + 2)),
                               s);
        launched("score_histogram");
    }
}

struct PrefilterImpl {
    float* P = nullptr;   // FP32 master parameters
    float* G = nullptr;   // gradients
    float* M = nullptr;   // Adam moments
    float* V = nullptr;
    __bf16* B = nullptr;  // BF16 operand copies
    uint64_t step = 0;

    ~PrefilterImpl() {
        for (void* p : {static_cast<void*>(P), static_cast<void*>(G), static_cast<void*>(M),
                        static_cast<void*>(V), static_cast<void*>(B)})
            if (p) (void)hipFree(p);
    }
};

void pf_check_set(uint32_t n, uint32_t max_len) {
    if (max_len > pf::kMaxLen)
        throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: the prefilter encoder takes sequences "
                                            "of at most " + std::to_string(pf::kMaxLen) +
                                                " residues (got " + std::to_string(max_len) + ")"};
    if (n == 0) throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: empty sequence set"};
}

// Pair indices past either set: ST_INVALID_ARGUMENT naming the first such pair.
bool bad_pair_index(st_status* status, const uint32_t* pq, const uint32_t* pr, uint64_t n_pairs,
                    uint64_t n_queries, uint64_t n_refs) {
    for (uint64_t t = 0; t < n_pairs; ++t)
        if (pq[t] >= n_queries || pr[t] >= n_refs) {
            set_status(status, ST_INVALID_ARGUMENT,
                       "seqtree_hip: pair %llu indexes past the query or reference set",
                       (unsigned long long)t);
            return true;
        }
    return false;
}

// Device copies and outputs of a host-memory pair traceback on a private stream: both sets,
// the pair list, the ops, the int64 scores (or, for the penalty form, the int32 scores) and the
// per-pair op count and first aligned residues.
struct PairTraceback {
    HostCall call;
    Arena ar;
    HostSet qh, rh;
    DeviceSet qd, rd;
    uint32_t *pq, *pr, *meta;
    uint8_t* ops;
    int64_t* score = nullptr;
    int32_t* score32 = nullptr;
    uint64_t n, stride;

    PairTraceback(const uint8_t* q, const uint64_t* q_off, uint64_t nq, const uint8_t* r,
                  const uint64_t* r_off, uint64_t nr, const uint32_t* pair_q,
                  const uint32_t* pair_r, uint64_t n_pairs, uint64_t ops_stride, bool penalty)
        : ar(call.stream()), qh(host_set(q, q_off, nq)), rh(host_set(r, r_off, nr)),
          qd((ar.prefer(call.scratch(), kContextScratch), upload(ar, qh, call))),
          rd(upload(ar, rh, call)), n(n_pairs), stride(ops_stride) {
        pq = ar.get<uint32_t>(n);
        pr = ar.get<uint32_t>(n);
        call.put(pq, pair_q, n);
        call.put(pr, pair_r, n);
        ops = ar.get<uint8_t>(n * stride);
        if (penalty) score32 = ar.get<int32_t>(n);
        else score = ar.get<int64_t>(n);
        meta = ar.get<uint32_t>(3 * n);
    }
    hipStream_t s() const { return call.stream(); }
    void fetch(uint8_t* ops_h, uint32_t* n_ops, uint32_t* q_begin, uint32_t* r_begin) {
        call.get(ops_h, ops, n * stride);
        call.get(n_ops, meta, n);
        if (q_begin) call.get(q_begin, meta + n, n);
        if (r_begin) call.get(r_begin, meta + 2 * n, n);
    }
};

// ---- allocation-free gap-block matrix (st_gapblock_matrix_device_noalloc) ------------------
// The device address of the static scratch array on the current device, looked up once per
// device with that device current and kept for the life of the process.
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;

void* noalloc_scratch() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    const auto it = g_scratch_addr.find(dev);
    if (it != g_scratch_addr.end()) return it->second;
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_addr.emplace(dev, p);
    return p;
}

// Marks the calling thread as inside an allocation-free call (no timing events).
struct NoAllocScope {
    NoAllocScope() { t_no_alloc = true; }
    ~NoAllocScope() { t_no_alloc = false; }
};

// An upper bound of the scratch a dense gap-block call carves (256-byte granules, as Arena):
// the penalty table, the prior cube with its padding, the packed references (words, lengths,
// columns) and the twin list with its sorted indices.
uint64_t gapblock_scratch_bound(uint64_t N, uint64_t K, uint32_t r_max, uint32_t A,
                                bool prior, uint32_t W) {
    auto granule = [](uint64_t bytes) { return (std::max<uint64_t>(bytes, 1) + 255) / 256 * 256; };
    const PackedRefs geo = packed_geometry(K, r_max);
    const uint64_t w1 = uint64_t(W) + 1;
    uint64_t b = granule(uint64_t(A) * A * 4);
    if (prior) b += granule((w1 * w1 * w1 + gb::kPriorPad) * 4);
    b += granule(uint64_t(geo.nwords + 1) * geo.npad * 4) + 2 * granule(uint64_t(geo.npad) * 4);
    b += granule((twin_capacity(N) + (N + 1) / 2) * sizeof(uint2));
    return b;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// single pairs: the scalar entry points
// ------------------------------------------------------------------------------------------
// A host-memory call with one pair (align_score, align_pair, Index::align, hamming,
// levenshtein) runs one single-pair kernel (single.hip) on a stream of its own and reads its
// results from pinned host memory: one launch and one synchronisation instead of the batch
// path's uploads, packing launches and downloads. The contexts (stream, pinned result) are
// kept for reuse, one per concurrent caller.
struct SingleContext {
    int device = 0;
    hipStream_t stream = nullptr;
    SingleOut* out = nullptr;
};
std::mutex g_single_mu;
std::vector<SingleContext*> g_single_free;

class SingleCall {
public:
    SingleCall() {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        {
            std::lock_guard<std::mutex> lk(g_single_mu);
            for (size_t i = 0; i < g_single_free.size(); ++i)
                if (g_single_free[i]->device == dev) {
                    c_ = g_single_free[i];
                    g_single_free.erase(g_single_free.begin() + i);
                    break;
                }
        }
        if (c_) return;
        auto* c = new SingleContext();
        c->device = dev;
        check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
        check(hipHostMalloc(reinterpret_cast<void**>(&c->out), sizeof(SingleOut),
                            hipHostMallocDefault),
              "hipHostMalloc");
        c_ = c;
    }
    ~SingleCall() {
        if (!done_) return;  // a failed call leaves its context behind
        std::lock_guard<std::mutex> lk(g_single_mu);
        g_single_free.push_back(c_);
    }
    SingleCall(const SingleCall&) = delete;
    SingleCall& operator=(const SingleCall&) = delete;

    hipStream_t stream() const { return c_->stream; }
    // Launches the kernel and waits for it; the results are valid until the call ends.
    const SingleOut& run(SingleKind kind, SingleArgs& a, const char* label) {
        a.out = c_->out;
        {
            LaunchTimer t(label, c_->stream);
            launch_single(kind, a, c_->stream);
        }
        launched(label);
        sync(c_->stream);
        done_ = true;
        return *c_->out;
    }

private:
    SingleContext* c_ = nullptr;
    bool done_ = false;
};

// The pair as single-kernel arguments when it fits them: 1 <= m <= max_q, 1 <= n <=
// kSingleMaxR (0 allowed for both when `empty_ok`), and every code below `codes_below` when
// that is nonzero (the kernels index their tables by code).
bool single_args(SingleArgs& a, const uint8_t* q, uint64_t m, const uint8_t* r, uint64_t n,
                 uint32_t max_q, uint32_t codes_below, bool empty_ok = false) {
    if (m > max_q || n > kSingleMaxR || (!empty_ok && (m == 0 || n == 0))) return false;
    if (codes_below) {
        for (uint64_t t = 0; t < m; ++t)
            if (q[t] >= codes_below) return false;
        for (uint64_t t = 0; t < n; ++t)
            if (r[t] >= codes_below) return false;
    }
    a = SingleArgs{};
    std::memcpy(a.q, q, m);
    std::memcpy(a.r, r, n);
    a.m = uint32_t(m);
    a.n = uint32_t(n);
    return true;
}

// Gotoh on int32 cells needs (m + n + 2) * max(|s|, go, ge) < 2^28 (the traceback kernels'
// bound, sentinel -2^29).
bool single_gotoh_range(const int32_t* table, uint32_t A, uint64_t m, uint64_t n, int32_t go,
                        int32_t ge) {
    if (go < 0 || ge < 0) return false;
    int64_t mx = std::max<int64_t>(go, ge);
    for (uint32_t e = 0; e < A * A; ++e) mx = std::max(mx, iabs64(table[e]));
    return (int64_t(m) + int64_t(n) + 2) * mx < (int64_t(1) << 28);
}

// One Gotoh pair (the query's codes and the reference's, similarity table sim_h) through the
// single-pair kernel: the score, or with `trace` the traceback too. False when it does not fit.
bool single_gotoh(const uint8_t* q, uint64_t m, const uint8_t* r, uint64_t n,
                  const int32_t* sim_h, uint32_t A, bool local, int32_t go, int32_t ge, bool trace,
                  const char* label, SingleOut* res) {
    SingleArgs a;
    if (!single_args(a, q, m, r, n, kSingleMaxQ, A) || !single_gotoh_range(sim_h, A, m, n, go, ge))
        return false;
    SingleCall call;
    Arena ar(call.stream());
    a.sim = table_on_device(ar, sim_h, size_t(A) * A, 0, call.stream()).dev;
    a.A = A;
    a.go = go;
    a.ge = ge;
    const SingleKind kind = trace ? (local ? SingleKind::TracebackLocal : SingleKind::TracebackGlobal)
                                  : (local ? SingleKind::GotohLocal : SingleKind::GotohGlobal);
    *res = call.run(kind, a, label);
    return true;
}

extern "C" {

const char* st_version(void) { return kVersion; }

int32_t st_device_alloc(uint64_t bytes, void** ptr, st_status* status) {
    *ptr = nullptr;
    return guarded(status, [&] {
        require_device();
        check(hipMalloc(ptr, std::max<uint64_t>(bytes, 1)), "hipMalloc");
        return ok(status);
    });
}

int32_t st_device_free(void* ptr) { return ptr && hipFree(ptr) != hipSuccess ? ST_DEVICE_ERROR : ST_OK; }

int32_t st_memcpy(void* dst, const void* src, uint64_t bytes, int32_t kind, st_status* status) {
    return guarded(status, [&] {
        const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                              : kind == 2 ? hipMemcpyDeviceToHost
                                          : hipMemcpyDeviceToDevice;
        if (bytes) check(hipMemcpy(dst, src, bytes, k), "hipMemcpy");
        return ok(status);
    });
}

int32_t st_synchronize(st_status* status) {
    return guarded(status, [&] {
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        return ok(status);
    });
}

int32_t st_probe_valu(int32_t kind, char* name, uint64_t name_capacity, double* lane_ops_per_s,
                      st_status* status) {
    return guarded(status, [&] {
        const char* nm = probe_name(kind);
        if (!nm) {
            set_status(status, ST_INVALID_ARGUMENT, "seqtree_hip: no VALU probe of kind %d", kind);
            return int(ST_INVALID_ARGUMENT);
        }
        if (name && name_capacity) std::snprintf(name, name_capacity, "%s", nm);
        require_device();
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
              "hipDeviceGetAttribute");
        const uint32_t blocks = uint32_t(cus) * 32, iters = 16384;
        uint32_t* out = nullptr;
        check(hipMalloc(&out, size_t(blocks) * 256 * sizeof(uint32_t)), "hipMalloc");
        hipEvent_t a = nullptr, b = nullptr;
        float best = 0.f;
        try {
            check(hipEventCreate(&a), "hipEventCreate");
            check(hipEventCreate(&b), "hipEventCreate");
            for (int rep = 0; rep < 4; ++rep) {  // the first run is a warm-up
                check(hipEventRecord(a, nullptr), "hipEventRecord");
                launch_probe(kind, out, blocks, iters, nullptr);
                check(hipGetLastError(), "probe launch");
                check(hipEventRecord(b, nullptr), "hipEventRecord");
                check(hipEventSynchronize(b), "hipEventSynchronize");
                float ms = 0.f;
                check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
                if (rep == 1 || (rep > 1 && ms < best)) best = ms;
            }
        } catch (...) {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            (void)hipFree(out);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        check(hipFree(out), "hipFree");
        *lane_ops_per_s = double(blocks) * 256.0 * 16.0 * iters / (double(best) * 1e-3);
        return ok(status);
    });
}


struct st_prefilter {
    PrefilterImpl impl;
};

uint64_t st_prefilter_param_count(void) { return pf::kParams; }

int32_t st_prefilter_create(uint64_t seed, st_prefilter** model, st_status* status) {
    *model = nullptr;
    return guarded(status, [&] {
        require_device();
        auto M = std::make_unique<st_prefilter>();
        PrefilterImpl& m = M->impl;
        for (float** p : {&m.P, &m.G, &m.M, &m.V})
            check(hipMalloc(reinterpret_cast<void**>(p), pf::kParams * sizeof(float)), "hipMalloc");
        check(hipMalloc(reinterpret_cast<void**>(&m.B), pf::kBf16Params * sizeof(__bf16)),
              "hipMalloc");
        // He-normal convolution weights, N(0, 1/2) embedding, zero biases
        std::vector<float> h(pf::kParams, 0.f);
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        for (uint32_t i = 0; i < pf::kEmbRows * pf::kEmbDim; ++i) h[pf::kOffEmb + i] = 0.5f * nd(rng);
        const uint32_t offs[3] = {pf::kOffW1, pf::kOffW2, pf::kOffW3};
        const uint32_t cin[3] = {pf::kEmbDim, pf::kChannels, pf::kChannels};
        for (int l = 0; l < 3; ++l) {
            const float sd = std::sqrt(2.f / float(5 * cin[l]));
            for (uint32_t i = 0; i < 5 * cin[l] * pf::kChannels; ++i) h[offs[l] + i] = sd * nd(rng);
        }
        check(hipMemcpy(m.P, h.data(), pf::kParams * sizeof(float), hipMemcpyHostToDevice), "H2D");
        check(hipMemset(m.M, 0, pf::kParams * sizeof(float)), "hipMemset");
        check(hipMemset(m.V, 0, pf::kParams * sizeof(float)), "hipMemset");
        pf::launch_pack_weights(m.P, m.B, nullptr);
        check(hipDeviceSynchronize(), "prefilter init");
        *model = M.release();
        return ok(status);
    });
}

void st_prefilter_destroy(st_prefilter* model) {
    if (model) (void)hipDeviceSynchronize();
    delete model;
}

int32_t st_prefilter_get_params(const st_prefilter* model, float* params, uint64_t count,
                                st_status* status) {
    return guarded(status, [&] {
        if (!model || count != pf::kParams)
            throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: parameter count mismatch"};
        check(hipDeviceSynchronize(), "sync");
        check(hipMemcpy(params, model->impl.P, count * sizeof(float), hipMemcpyDeviceToHost), "D2H");
        return ok(status);
    });
}

int32_t st_prefilter_set_params(st_prefilter* model, const float* params, uint64_t count,
                                st_status* status) {
    return guarded(status, [&] {
        if (!model || count != pf::kParams)
            throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: parameter count mismatch"};
        check(hipDeviceSynchronize(), "sync");
        check(hipMemcpy(model->impl.P, params, count * sizeof(float), hipMemcpyHostToDevice), "H2D");
        pf::launch_pack_weights(model->impl.P, model->impl.B, nullptr);
        check(hipDeviceSynchronize(), "prefilter pack");
        return ok(status);
    });
}

int32_t st_prefilter_get_grads(const st_prefilter* model, float* grads, uint64_t count,
                               st_status* status) {
    return guarded(status, [&] {
        if (!model || count != pf::kParams)
            throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: parameter count mismatch"};
        check(hipDeviceSynchronize(), "sync");
        check(hipMemcpy(grads, model->impl.G, count * sizeof(float), hipMemcpyDeviceToHost), "D2H");
        return ok(status);
    });
}

int32_t st_prefilter_pair_distance(const void* q_emb, const float* q_norm, const void* r_emb,
                                   const float* r_norm, const uint32_t* pair_query,
                                   const uint32_t* pair_ref, uint64_t n_pairs, float* distance,
                                   void* stream, st_status* status) {
    return guarded(status, [&] {
        require_device();
        if (n_pairs >= (uint64_t(1) << 33))
            throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: too many pairs for one call"};
        const hipStream_t s = static_cast<hipStream_t>(stream);
        {
            LaunchTimer t("prefilter_pair_distance", s);
            pf::launch_pair_distance(static_cast<const __bf16*>(q_emb), q_norm,
                                     static_cast<const __bf16*>(r_emb), r_norm, pair_query,
                                     pair_ref, n_pairs, distance, s);
        }
        launched("prefilter_pair_distance");
        return ok(status);
    });
}

int32_t st_select_le(const int32_t* values, uint64_t n, int32_t cutoff, uint64_t capacity,
                     uint32_t* index, uint64_t* count, void* stream, st_status* status) {
    return st_select_pairs_le(values, nullptr, nullptr, n, cutoff, capacity, index, nullptr,
                              nullptr, count, stream, status);
}

int32_t st_select_pairs_le(const int32_t* values, const uint32_t* pair_a, const uint32_t* pair_b,
                           uint64_t n, int32_t cutoff, uint64_t capacity, uint32_t* out_a,
                           uint32_t* out_b, int32_t* out_value, uint64_t* count, void* stream,
                           st_status* status) {
    return guarded(status, [&] {
        require_device();
        if ((pair_a == nullptr) != (pair_b == nullptr) ||
            (pair_a && (out_b == nullptr || out_value == nullptr)))
            throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: incomplete pair arrays"};
        const hipStream_t s = static_cast<hipStream_t>(stream);
        Arena ar(s);
        unsigned long long* c = ar.get<unsigned long long>(1);
        launch_fill_u32(reinterpret_cast<uint32_t*>(c), 2, 0u, s);
        {
            LaunchTimer t("select_le", s);
            pf::launch_select_le(values, n, cutoff, pair_a, pair_b, out_a, out_b, out_value,
                                 capacity, c, s);
        }
        launched("select_le");
        unsigned long long h = 0;
        d2h(&h, c, 1, s);
        sync(s);
        *count = h;
        return ok(status);
    });
}

int32_t st_prefilter_train_step(st_prefilter* model, const uint8_t* codes,
                                const uint64_t* offsets, uint32_t n_seqs, uint32_t max_len,
                                const uint32_t* pair_a, const uint32_t* pair_b,
                                const float* target, const float* weight, uint32_t n_pairs,
                                float lr, float* loss_out, void* stream, st_status* status) {
    return guarded(status, [&] {
        if (!model) throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: null prefilter model"};
        pf_check_set(n_seqs, max_len);
        require_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        PrefilterImpl& m = model->impl;
        const PfRows rows = pf_rows(read_offset(offsets, 0, s), read_offset(offsets, n_seqs, s),
                                    n_seqs);
        const uint64_t R = rows.R;
        Arena ar(s);
        const PfActs x0 = pf_acts(ar, R, pf::kEmbDim, s), h1 = pf_acts(ar, R, pf::kChannels, s),
                     h2 = pf_acts(ar, R, pf::kChannels, s), h3 = pf_acts(ar, R, pf::kChannels, s),
                     ga = pf_acts(ar, R, pf::kChannels, s), gb = pf_acts(ar, R, pf::kChannels, s),
                     dx0 = pf_acts(ar, R, pf::kEmbDim, s);
        uint8_t* valid = ar.get<uint8_t>(R);
        float* ef = ar.get<float>(uint64_t(n_seqs) * pf::kChannels);
        float* dE = ar.get<float>(uint64_t(n_seqs) * pf::kChannels);
        float* loss = ar.get<float>(1);
        launch_fill_u32(reinterpret_cast<uint32_t*>(dE), uint64_t(n_seqs) * pf::kChannels, 0u, s);
        launch_fill_u32(reinterpret_cast<uint32_t*>(loss), 1, 0u, s);
        launch_fill_u32(reinterpret_cast<uint32_t*>(m.G), pf::kParams, 0u, s);

        pf_forward(m.B, m.P, codes, offsets, n_seqs, rows, x0, h1, h2, h3, valid, s);
        {
            LaunchTimer t("prefilter_train", s);
            pf::launch_pool(h3.rows, offsets, n_seqs, ef, nullptr, nullptr, s);
            pf::launch_pair_loss(ef, pair_a, pair_b, target, weight, n_pairs, dE, loss, s);
            // backward: layer 3, 2, 1, embedding (rows past rows_used stay zero)
            launch_fill_u32(reinterpret_cast<uint32_t*>(ga.rows + rows.used * pf::kChannels),
                            (R - rows.used) * pf::kChannels / 2, 0u, s);
            pf::launch_pool_bwd(h3.rows, dE, offsets, n_seqs, ga.rows, s);
            pf::launch_dw(3, h2.rows, ga.rows, R, m.G + pf::kOffW3, m.G + pf::kOffB3, s);
            pf::launch_conv(3, 1, ga.rows, m.B + pf::kOffWf3, nullptr, h2.rows, valid, gb.rows, R,
                            s);
            pf::launch_dw(2, h1.rows, gb.rows, R, m.G + pf::kOffW2, m.G + pf::kOffB2, s);
            pf::launch_conv(2, 1, gb.rows, m.B + pf::kOffWf2, nullptr, h1.rows, valid, ga.rows, R,
                            s);
            pf::launch_dw(1, x0.rows, ga.rows, R, m.G + pf::kOffW1, m.G + pf::kOffB1, s);
            pf::launch_conv(1, 2, ga.rows, m.B + pf::kOffWf1, nullptr, nullptr, valid, dx0.rows, R,
                            s);
            pf::launch_embed_bwd(dx0.rows, codes, offsets, n_seqs, m.G + pf::kOffEmb, s);
            m.step += 1;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, float(m.step)), c2 = 1.f - std::pow(b2, float(m.step));
            pf::launch_adam(m.P, m.G, m.M, m.V, pf::kParams, lr, b1, b2, c1, c2, 1e-8f, s);
            pf::launch_pack_weights(m.P, m.B, s);
        }
        launched("prefilter train");
        float h = 0.f;
        d2h(&h, loss, 1, s);
        sync(s);
        if (loss_out) *loss_out = n_pairs ? h / float(n_pairs) : 0.f;
        return ok(status);
    });
}

int32_t st_prefilter_encode(const st_prefilter* model, const uint8_t* codes,
                            const uint64_t* offsets, uint32_t n, uint32_t max_len, void* emb,
                            float* norm, void* stream, st_status* status) {
    return guarded(status, [&] {
        if (!model) throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: null prefilter model"};
        pf_check_set(n, max_len);
        require_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const PrefilterImpl& m = model->impl;
        // blocks of about 2^22 rows (256 MB of X0)
        const uint32_t blk = uint32_t(std::min<uint64_t>(n, std::max<uint64_t>(1024, (uint64_t(1) << 22) / (max_len + 2))));
        const uint64_t Rb = (uint64_t(blk) * (max_len + 2) + 127) / 128 * 128;
        // the offsets at the block boundaries (blocks of blk sequences, then n)
        const uint32_t nb = (n + blk - 1) / blk;
        Arena ar(s);
        uint64_t* dbound = ar.get<uint64_t>(nb + 1);
        pf::launch_gather_offsets(offsets, blk, n, nb + 1, dbound, s);
        std::vector<uint64_t> bound(nb + 1);
        d2h(bound.data(), dbound, nb + 1, s);
        sync(s);
        const PfActs x0 = pf_acts(ar, Rb, pf::kEmbDim, s, kEncX0Guard);
        const PfEncodeScratch sc = pf_encode_scratch(ar, Rb);
        uint8_t* valid = ar.get<uint8_t>(Rb + kEncValidHead + kEncValidTail) + kEncValidHead;
        __bf16* out = static_cast<__bf16*>(emb);
        for (uint32_t s0 = 0; s0 < n; s0 += blk) {
            const uint32_t cnt = std::min(blk, n - s0);
            const PfRows rows = pf_rows(bound[s0 / blk], bound[s0 / blk + 1], cnt);
            if (rows.R > Rb)
                throw HipError{ST_INVALID_ARGUMENT,
                               "seqtree_hip: a sequence is longer than max_len"};
            pf_encode_forward(m.B, m.P, codes, offsets + s0, cnt, rows, x0, valid, sc,
                              out + uint64_t(s0) * pf::kChannels, norm + s0, s);
        }
        launched("prefilter encode");
        return ok(status);
    });
}

int32_t st_prefilter_screen(const void* q_emb, const float* q_norm, uint32_t n_queries,
                            const void* r_emb, const float* r_norm, uint32_t n_refs, float tau,
                            uint64_t capacity, uint32_t* hit_query, uint32_t* hit_ref,
                            uint64_t* hit_count, void* stream, st_status* status) {
    return guarded(status, [&] {
        require_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        int dev = 0, cus = 256;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attr");
        Arena ar(s);
        unsigned long long* count = ar.get<unsigned long long>(1);
        launch_fill_u32(reinterpret_cast<uint32_t*>(count), 2, 0u, s);
        {
            LaunchTimer t("prefilter_screen", s);
            pf::launch_gram(static_cast<const __bf16*>(q_emb), q_norm, n_queries,
                            static_cast<const __bf16*>(r_emb), r_norm, n_refs, tau, hit_query,
                            hit_ref, capacity, count, uint32_t(cus), s);
        }
        launched("prefilter screen");
        unsigned long long h = 0;
        d2h(&h, count, 1, s);
        sync(s);
        *hit_count = h;
        return ok(status);
    });
}

int32_t st_probe_mfma(double* flops_per_s, st_status* status) {
    return guarded(status, [&] {
        require_device();
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
              "hipDeviceGetAttribute");
        const uint32_t blocks = uint32_t(cus) * 16, iters = 2000;
        float* out = nullptr;
        check(hipMalloc(&out, size_t(blocks) * 256 * sizeof(float)), "hipMalloc");
        hipEvent_t a = nullptr, b = nullptr;
        float best = 0.f;
        double flops = 0.0;
        try {
            check(hipEventCreate(&a), "hipEventCreate");
            check(hipEventCreate(&b), "hipEventCreate");
            for (int rep = 0; rep < 4; ++rep) {  // the first run is a warm-up
                check(hipEventRecord(a, nullptr), "hipEventRecord");
                flops = launch_probe_mfma(out, blocks, iters, nullptr);
                check(hipGetLastError(), "probe launch");
                check(hipEventRecord(b, nullptr), "hipEventRecord");
                check(hipEventSynchronize(b), "hipEventSynchronize");
                float ms = 0.f;
                check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
                if (rep == 1 || (rep > 1 && ms < best)) best = ms;
            }
        } catch (...) {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            (void)hipFree(out);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        check(hipFree(out), "hipFree");
        *flops_per_s = flops / (double(best) * 1e-3);
        return ok(status);
    });
}

int32_t st_probe_hbm_write(uint64_t bytes, double* bytes_per_s, st_status* status) {
    return guarded(status, [&] {
        require_device();
        if (bytes < (uint64_t(1) << 20) || bytes % 16)
            throw HipError{ST_INVALID_ARGUMENT,
                           "seqtree_hip: the write probe takes at least 1 MiB, a multiple of 16"};
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
              "hipDeviceGetAttribute");
        void* buf = nullptr;
        check(hipMalloc(&buf, bytes), "hipMalloc");
        hipEvent_t a = nullptr, b = nullptr;
        double best = 0.0;
        try {
            check(hipEventCreate(&a), "hipEventCreate");
            check(hipEventCreate(&b), "hipEventCreate");
            // grids of 4 and 64 workgroups per CU; the first run of each is a warm-up
            for (uint32_t per_cu : {4u, 64u})
                for (int rep = 0; rep < 3; ++rep) {
                    check(hipEventRecord(a, nullptr), "hipEventRecord");
                    launch_probe_fill(buf, bytes, uint32_t(cus) * per_cu, uint32_t(rep), nullptr);
                    check(hipGetLastError(), "probe launch");
                    check(hipEventRecord(b, nullptr), "hipEventRecord");
                    check(hipEventSynchronize(b), "hipEventSynchronize");
                    float ms = 0.f;
                    check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
                    if (rep > 0) best = std::max(best, double(bytes) / (double(ms) * 1e-3));
                }
        } catch (...) {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            (void)hipFree(buf);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        check(hipFree(buf), "hipFree");
        *bytes_per_s = best;
        return ok(status);
    });
}

int32_t st_probe_lds(int32_t dwords, double* bytes_per_s, st_status* status) {
    return guarded(status, [&] {
        if (dwords != 1 && dwords != 2 && dwords != 4)
            throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: the LDS probe reads 1, 2 or 4 dwords"};
        require_device();
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
              "hipDeviceGetAttribute");
        const uint32_t blocks = uint32_t(cus) * 16, iters = 4096 / uint32_t(dwords);
        uint32_t* out = nullptr;
        check(hipMalloc(&out, size_t(blocks) * 256 * sizeof(uint32_t)), "hipMalloc");
        hipEvent_t a = nullptr, b = nullptr;
        double best = 0.0;
        try {
            check(hipEventCreate(&a), "hipEventCreate");
            check(hipEventCreate(&b), "hipEventCreate");
            for (int rep = 0; rep < 4; ++rep) {  // the first run is a warm-up
                check(hipEventRecord(a, nullptr), "hipEventRecord");
                const double bytes = launch_probe_lds(dwords, out, blocks, iters, nullptr);
                check(hipGetLastError(), "probe launch");
                check(hipEventRecord(b, nullptr), "hipEventRecord");
                check(hipEventSynchronize(b), "hipEventSynchronize");
                float ms = 0.f;
                check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
                if (rep > 0) best = std::max(best, bytes / (double(ms) * 1e-3));
            }
        } catch (...) {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            (void)hipFree(out);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        check(hipFree(out), "hipFree");
        *bytes_per_s = best;
        return ok(status);
    });
}

void st_timing_enable(int32_t on) {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing = on != 0;
}

int32_t st_timing_collect(char* report, uint64_t capacity) {
    std::vector<Pending> done;
    {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        done.swap(g_pending);
    }
    std::map<std::string, std::pair<double, int>> acc;
    int bad = 0;
    for (auto& p : done) {
        float ms = 0.f;
        if (hipEventSynchronize(p.stop) == hipSuccess &&
            hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            auto& e = acc[p.label];
            e.first += ms;
            e.second += 1;
        } else {
            ++bad;
        }
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    std::string out;
    for (auto& kv : acc) {
        char line[256];
        std::snprintf(line, sizeof(line), "%s %d %.6f\n", kv.first.c_str(), kv.second.second,
                      kv.second.first);
        out += line;
    }
    if (report && capacity) {
        std::strncpy(report, out.c_str(), capacity - 1);
        report[capacity - 1] = 0;
    }
    return bad ? -bad : int32_t(acc.size());
}

int32_t st_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int32_t st_host_alloc(uint64_t bytes, void** ptr, st_status* status) {
    if (!ptr) {
        set_status(status, ST_INVALID_ARGUMENT, "seqtree_hip: st_host_alloc needs a pointer");
        return ST_INVALID_ARGUMENT;
    }
    *ptr = nullptr;
    void* p = nullptr;
    if (bytes == 0 || hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();  // leave no error behind for the next call
        set_status(status, ST_OUT_OF_MEMORY, "seqtree_hip: cannot pin %llu bytes of host memory",
                   static_cast<unsigned long long>(bytes));
        return ST_OUT_OF_MEMORY;
    }
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        g_pinned[reinterpret_cast<uintptr_t>(p)] = bytes;
    }
    *ptr = p;
    return ok(status);
}

void st_host_free(void* ptr) {
    if (!ptr) return;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        if (g_pinned.erase(reinterpret_cast<uintptr_t>(ptr)) == 0) return;
    }
    (void)hipHostFree(ptr);
}

int32_t st_gapblock_matrix_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                  uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                                  const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                                  const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                                  int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                                  int32_t* out, void* stream, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    return guarded(status, [&] {
        require_device();
        gapblock_device(q_codes, q_offsets, n_queries, q_max_len, r_codes, r_offsets, n_refs,
                        r_max_len, pen, alphabet_size, gap_open, gap_extend, prior, prior_width,
                        out, static_cast<hipStream_t>(stream));
        return ok(status);
    });
}

int32_t st_gapblock_matrix_device_noalloc(const uint8_t* q_codes, const uint64_t* q_offsets,
                                          uint64_t n_queries, uint32_t q_max_len,
                                          const uint8_t* r_codes, const uint64_t* r_offsets,
                                          uint64_t n_refs, uint32_t r_max_len, const int32_t* pen,
                                          uint32_t alphabet_size, int32_t gap_open,
                                          int32_t gap_extend, const int32_t* prior,
                                          uint32_t prior_width, int32_t* out, void* stream,
                                          st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    return guarded(status, [&] {
        require_device();
        NoAllocScope no_alloc;
        if (n_queries == 0 || n_refs == 0) return ok(status);
        const uint64_t need = gapblock_scratch_bound(n_queries, n_refs, r_max_len, alphabet_size,
                                                     prior != nullptr, prior_width);
        if (need > kNoAllocScratchBytes)
            throw HipError{ST_INVALID_ARGUMENT,
                           "seqtree_hip: st_gapblock_matrix_device_noalloc needs " +
                               std::to_string(need) + " bytes of scratch for these sets; its "
                               "static scratch holds " + std::to_string(kNoAllocScratchBytes)};
        gapblock_device(q_codes, q_offsets, n_queries, q_max_len, r_codes, r_offsets, n_refs,
                        r_max_len, pen, alphabet_size, gap_open, gap_extend, prior, prior_width,
                        out, static_cast<hipStream_t>(stream), noalloc_scratch(),
                        kNoAllocScratchBytes);
        return ok(status);
    });
}

int32_t st_gapblock_matrix(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                           const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                           const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                           int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                           int32_t* out, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    if (n_queries == 0 || n_refs == 0) return ok(status);
    return guarded(status, [&] {
        require_device();
        const HostSet qh = host_set(q_codes, q_offsets, n_queries);
        const HostSet rh = host_set(r_codes, r_offsets, n_refs);
        HostCall call;
        const hipStream_t s = call.stream();
        Arena ar(s);
        ar.prefer(call.scratch(), kContextScratch);
        const DeviceSet qd = upload(ar, qh, call);
        const DeviceSet rd = upload(ar, rh, call);
        run_rows(qh, n_refs, out, call, ar, qd,
                 [&](const uint64_t* qoff, uint64_t, uint64_t rows, int32_t* dout) {
                     gapblock_device(qd.bytes, qoff, rows, qh.max_len, rd.bytes, rd.off, n_refs,
                                     rh.max_len, pen, alphabet_size, gap_open, gap_extend, prior,
                                     prior_width, dout, s);
                 });
        return ok(status);
    });
}

int32_t st_gapblock_histogram_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                     uint64_t n_queries, uint32_t q_max_len,
                                     const uint8_t* r_codes, const uint64_t* r_offsets,
                                     uint64_t n_refs, uint32_t r_max_len, const int32_t* pen,
                                     uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                                     const int32_t* prior, uint32_t prior_width, uint32_t bins,
                                     uint64_t* histogram, void* stream, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    return guarded(status, [&] {
        require_device();
        gapblock_histogram_device(q_codes, q_offsets, n_queries, q_max_len, r_codes, r_offsets,
                                  n_refs, r_max_len, pen, alphabet_size, gap_open, gap_extend,
                                  prior, prior_width, bins, histogram,
                                  static_cast<hipStream_t>(stream));
        return ok(status);
    });
}

int32_t st_gapblock_histogram(const uint8_t* q_codes, const uint64_t* q_offsets,
                              uint64_t n_queries, const uint8_t* r_codes,
                              const uint64_t* r_offsets, uint64_t n_refs, const int32_t* pen,
                              uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                              const int32_t* prior, uint32_t prior_width, uint32_t bins,
                              uint64_t* histogram, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    return guarded(status, [&] {
        require_device();
        if (n_queries == 0) return ok(status);
        const HostSet qh = host_set(q_codes, q_offsets, n_queries);
        const HostSet rh = host_set(r_codes, r_offsets, n_refs);
        HostCall call;
        const hipStream_t s = call.stream();
        Arena ar(s);
        ar.prefer(call.scratch(), kContextScratch);
        const DeviceSet qd = upload(ar, qh, call);
        const DeviceSet rd = upload(ar, rh, call);
        uint64_t* d = ar.get<uint64_t>(n_queries * (uint64_t(bins) + 2));
        gapblock_histogram_device(qd.bytes, qd.off, n_queries, qh.max_len, rd.bytes, rd.off,
                                  n_refs, rh.max_len, pen, alphabet_size, gap_open, gap_extend,
                                  prior, prior_width, bins, d, s);
        call.get(histogram, d, n_queries * (uint64_t(bins) + 2));
        return ok(status);
    });
}

int32_t st_gapblock_threshold_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                     uint64_t n_queries, uint32_t q_max_len,
                                     const uint8_t* r_codes, const uint64_t* r_offsets,
                                     uint64_t n_refs, uint32_t r_max_len, const int32_t* pen,
                                     uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                                     const int32_t* prior, uint32_t prior_width, int32_t cutoff,
                                     uint64_t capacity, uint32_t* hit_query, uint32_t* hit_ref,
                                     int32_t* hit_score, uint64_t* hit_count, void* stream,
                                     st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    return guarded(status, [&] {
        require_device();
        GapblockHits h{cutoff, capacity, hit_query, hit_ref, hit_score,
                       reinterpret_cast<unsigned long long*>(hit_count)};
        gapblock_threshold_device(q_codes, q_offsets, n_queries, q_max_len, r_codes, r_offsets,
                                  n_refs, r_max_len, pen, alphabet_size, gap_open, gap_extend,
                                  prior, prior_width, h, static_cast<hipStream_t>(stream));
        return ok(status);
    });
}

int32_t st_gapblock_threshold(const uint8_t* q_codes, const uint64_t* q_offsets,
                              uint64_t n_queries, const uint8_t* r_codes,
                              const uint64_t* r_offsets, uint64_t n_refs, const int32_t* pen,
                              uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                              const int32_t* prior, uint32_t prior_width, int32_t cutoff,
                              uint64_t capacity, uint32_t* hit_query, uint32_t* hit_ref,
                              int32_t* hit_score, uint64_t* hit_count, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    *hit_count = 0;
    if (n_queries == 0 || n_refs == 0) return ok(status);
    return guarded(status, [&] {
        require_device();
        const HostSet qh = host_set(q_codes, q_offsets, n_queries);
        const HostSet rh = host_set(r_codes, r_offsets, n_refs);
        HostCall call;
        const hipStream_t s = call.stream();
        Arena ar(s);
        ar.prefer(call.scratch(), kContextScratch);
        const DeviceSet qd = upload(ar, qh, call);
        const DeviceSet rd = upload(ar, rh, call);
        const uint64_t cap = std::max<uint64_t>(capacity, 1);
        GapblockHits h{cutoff, capacity, ar.get<uint32_t>(cap), ar.get<uint32_t>(cap),
                       ar.get<int32_t>(cap), ar.get<unsigned long long>(1)};
        gapblock_threshold_device(qd.bytes, qd.off, n_queries, qh.max_len, rd.bytes, rd.off,
                                  n_refs, rh.max_len, pen, alphabet_size, gap_open, gap_extend,
                                  prior, prior_width, h, s);
        unsigned long long count = 0;
        call.get(&count, h.count, 1);
        const uint64_t kept = std::min<uint64_t>(count, capacity);
        call.get(hit_query, h.query, kept);
        call.get(hit_ref, h.ref, kept);
        call.get(hit_score, h.score, kept);
        *hit_count = count;
        return ok(status);
    });
}

int32_t st_gapblock_pairs_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                 uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                                 const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                                 const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                                 int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                                 const uint32_t* pair_query, const uint32_t* pair_ref,
                                 uint64_t n_pairs, int32_t* out, void* stream,
                                 st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    return guarded(status, [&] {
        require_device();
        gapblock_pairs_device(q_codes, q_offsets, n_queries, q_max_len, r_codes, r_offsets,
                              n_refs, r_max_len, pen, alphabet_size, gap_open, gap_extend, prior,
                              prior_width, pair_query, pair_ref, n_pairs, out,
                              static_cast<hipStream_t>(stream));
        return ok(status);
    });
}

int32_t st_gapblock_pairs(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                          const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                          const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                          int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                          const uint32_t* pair_query, const uint32_t* pair_ref, uint64_t n_pairs,
                          int32_t* out, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    if (n_pairs == 0) return ok(status);
    for (uint64_t t = 0; t < n_pairs; ++t)
        if (pair_query[t] >= n_queries || pair_ref[t] >= n_refs) {
            set_status(status, ST_INVALID_ARGUMENT,
                       "seqtree_hip: pair %llu indexes past the query or reference set",
                       (unsigned long long)t);
            return int(ST_INVALID_ARGUMENT);
        }
    return guarded(status, [&] {
        require_device();
        const HostSet qh = host_set(q_codes, q_offsets, n_queries);
        const HostSet rh = host_set(r_codes, r_offsets, n_refs);
        HostCall call;
        const hipStream_t s = call.stream();
        Arena ar(s);
        ar.prefer(call.scratch(), kContextScratch);
        const DeviceSet qd = upload(ar, qh, call);
        const DeviceSet rd = upload(ar, rh, call);
        uint32_t* pq = ar.get<uint32_t>(n_pairs);
        uint32_t* pr = ar.get<uint32_t>(n_pairs);
        int32_t* dout = ar.get<int32_t>(n_pairs);
        call.put(pq, pair_query, n_pairs);
        call.put(pr, pair_ref, n_pairs);
        gapblock_pairs_device(qd.bytes, qd.off, n_queries, qh.max_len, rd.bytes, rd.off, n_refs,
                              rh.max_len, pen, alphabet_size, gap_open, gap_extend, prior,
                              prior_width, pq, pr, n_pairs, dout, s);
        call.get(out, dout, n_pairs);
        return ok(status);
    });
}

int32_t st_align_pairs_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                              uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                              const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                              const uint32_t* pair_query, const uint32_t* pair_ref,
                              uint64_t n_pairs, const int32_t* sim, uint32_t alphabet_size,
                              int32_t local, int32_t gap_open, int32_t gap_extend, uint8_t* ops,
                              uint64_t ops_stride, int64_t* score, uint32_t* n_ops,
                              uint32_t* q_begin, uint32_t* r_begin, void* stream,
                              st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    (void)n_queries;
    (void)n_refs;
    return guarded(status, [&] {
        require_device();
        traceback_device(q_codes, q_offsets, q_max_len, r_codes, r_offsets, r_max_len,
                         pair_query, pair_ref, n_pairs, sim, alphabet_size, local, gap_open,
                         gap_extend, ops, ops_stride, score, n_ops, q_begin, r_begin,
                         static_cast<hipStream_t>(stream));
        return ok(status);
    });
}

int32_t st_align_pairs(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                       const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                       const uint32_t* pair_query, const uint32_t* pair_ref, uint64_t n_pairs,
                       const int32_t* sim, uint32_t alphabet_size, int32_t local,
                       int32_t gap_open, int32_t gap_extend, uint8_t* ops, uint64_t ops_stride,
                       int64_t* score, uint32_t* n_ops, uint32_t* q_begin, uint32_t* r_begin,
                       st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    if (n_pairs == 0) return ok(status);
    if (bad_pair_index(status, pair_query, pair_ref, n_pairs, n_queries, n_refs))
        return int(ST_INVALID_ARGUMENT);
    return guarded(status, [&] {
        require_device();
        if (n_pairs == 1) {  // align_pair: one pair
            const uint64_t q0 = q_offsets[pair_query[0]], r0 = r_offsets[pair_ref[0]];
            const uint64_t m = q_offsets[pair_query[0] + 1] - q0;
            const uint64_t n = r_offsets[pair_ref[0] + 1] - r0;
            SingleOut res;
            if (ops_stride >= m + n &&
                single_gotoh(q_codes + q0, m, r_codes + r0, n, sim, alphabet_size, local != 0,
                             gap_open, gap_extend, true,
                             local ? "single_traceback_local" : "single_traceback_global",
                             &res)) {
                std::memcpy(ops, res.ops, res.n_ops);
                score[0] = res.score;
                n_ops[0] = res.n_ops;
                q_begin[0] = res.q_begin;
                r_begin[0] = res.r_begin;
                return ok(status);
            }
        }
        PairTraceback t(q_codes, q_offsets, n_queries, r_codes, r_offsets, n_refs, pair_query,
                        pair_ref, n_pairs, ops_stride, false);
        traceback_device(t.qd.bytes, t.qd.off, t.qh.max_len, t.rd.bytes, t.rd.off, t.rh.max_len,
                         t.pq, t.pr, n_pairs, sim, alphabet_size, local, gap_open, gap_extend,
                         t.ops, ops_stride, t.score, t.meta, t.meta + n_pairs,
                         t.meta + 2 * n_pairs, t.s());
        t.fetch(ops, n_ops, q_begin, r_begin);
        t.call.get(score, t.score, n_pairs);
        return ok(status);
    });
}

int32_t st_index_align_pairs_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                    uint64_t n_queries, uint32_t q_max_len,
                                    const uint8_t* r_codes, const uint64_t* r_offsets,
                                    uint64_t n_refs, uint32_t r_max_len,
                                    const uint32_t* pair_query, const uint32_t* pair_ref,
                                    uint64_t n_pairs, const int32_t* pen, uint32_t alphabet_size,
                                    int32_t gap_open, int32_t gap_extend, uint8_t* ops,
                                    uint64_t ops_stride, int32_t* score, uint32_t* n_ops,
                                    void* stream, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    (void)n_queries;
    (void)n_refs;
    return guarded(status, [&] {
        require_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        Arena ar(s);
        // the first aligned residues are (0, 0) in a global alignment: device scratch only
        uint32_t* begins = ar.get<uint32_t>(2 * std::max<uint64_t>(n_pairs, 1));
        const std::vector<int32_t> neg = negated_penalties(pen, alphabet_size);
        traceback_device(q_codes, q_offsets, q_max_len, r_codes, r_offsets, r_max_len,
                         pair_query, pair_ref, n_pairs, neg.data(), alphabet_size, 0, gap_open,
                         gap_extend, ops, ops_stride, nullptr, n_ops, begins, begins + n_pairs,
                         s, score);
        return ok(status);
    });
}

int32_t st_index_align_pairs(const uint8_t* q_codes, const uint64_t* q_offsets,
                             uint64_t n_queries, const uint8_t* r_codes,
                             const uint64_t* r_offsets, uint64_t n_refs,
                             const uint32_t* pair_query, const uint32_t* pair_ref,
                             uint64_t n_pairs, const int32_t* pen, uint32_t alphabet_size,
                             int32_t gap_open, int32_t gap_extend, uint8_t* ops,
                             uint64_t ops_stride, int32_t* score, uint32_t* n_ops,
                             st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    if (n_pairs == 0) return ok(status);
    if (bad_pair_index(status, pair_query, pair_ref, n_pairs, n_queries, n_refs))
        return int(ST_INVALID_ARGUMENT);
    return guarded(status, [&] {
        require_device();
        const std::vector<int32_t> neg = negated_penalties(pen, alphabet_size);
        if (n_pairs == 1) {  // Index::align: one pair
            const uint64_t q0 = q_offsets[pair_query[0]], r0 = r_offsets[pair_ref[0]];
            const uint64_t m = q_offsets[pair_query[0] + 1] - q0;
            const uint64_t n = r_offsets[pair_ref[0] + 1] - r0;
            SingleOut res;
            if (ops_stride >= m + n &&
                single_gotoh(q_codes + q0, m, r_codes + r0, n, neg.data(), alphabet_size, false,
                             gap_open, gap_extend, true, "single_traceback_penalty", &res)) {
                std::memcpy(ops, res.ops, res.n_ops);
                score[0] = int32_t(-res.score);
                n_ops[0] = res.n_ops;
                return ok(status);
            }
        }
        PairTraceback t(q_codes, q_offsets, n_queries, r_codes, r_offsets, n_refs, pair_query,
                        pair_ref, n_pairs, ops_stride, true);
        traceback_device(t.qd.bytes, t.qd.off, t.qh.max_len, t.rd.bytes, t.rd.off, t.rh.max_len,
                         t.pq, t.pr, n_pairs, neg.data(), alphabet_size, 0, gap_open, gap_extend,
                         t.ops, ops_stride, nullptr, t.meta, t.meta + n_pairs,
                         t.meta + 2 * n_pairs, t.s(), t.score32);
        t.fetch(ops, n_ops, nullptr, nullptr);
        t.call.get(score, t.score32, n_pairs);
        return ok(status);
    });
}

struct st_batch_plan {
    BatchPlanImpl impl;
};

int32_t st_gapblock_batch_plan_create(const st_gapblock_job* jobs, uint32_t n_jobs,
                                      int32_t autotune, void* stream, st_batch_plan** plan,
                                      st_status* status) {
    *plan = nullptr;
    for (uint32_t t = 0; t < n_jobs; ++t)
        if (jobs[t].alphabet_size == 0 || jobs[t].alphabet_size > kMaxAlphabet)
            return bad_alphabet(status, jobs[t].alphabet_size);
    return guarded(status, [&] {
        require_device();
        const hipStream_t s = static_cast<hipStream_t>(stream);
        auto P = std::make_unique<st_batch_plan>();
        BatchPlanImpl& B = P->impl;
        B.jobs.assign(jobs, jobs + n_jobs);
        int dev = 0;
        hipDeviceProp_t prop;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        B.cus = prop.multiProcessorCount;

        // heaviest jobs first: a unit costs queries_per_unit x n_refs pairs
        B.order.resize(n_jobs);
        for (uint32_t t = 0; t < n_jobs; ++t) B.order[t] = t;
        std::stable_sort(B.order.begin(), B.order.end(), [&](uint32_t x, uint32_t y) {
            return jobs[x].n_refs > jobs[y].n_refs;
        });
        size_t lds = sizeof(int32_t);
        for (uint32_t t : B.order) {
            const st_gapblock_job& j = jobs[t];
            if (j.n_queries == 0 || j.n_refs == 0) continue;
            const uint32_t A = j.alphabet_size;
            int32_t* pen = B.alloc<int32_t>(size_t(A) * A);
            h2d(pen, j.pen, size_t(A) * A, s);
            if (j.prior && j.prior_width < std::max(j.q_max_len, j.r_max_len))
                throw HipError{ST_INVALID_ARGUMENT,
                               "seqtree_hip: batch job " + std::to_string(t) + ": prior_width " +
                                   std::to_string(j.prior_width) +
                                   " is below the longest sequence"};
            const size_t w1 = size_t(j.prior_width) + 1;
            int32_t* prior = nullptr;
            int64_t prior_abs = 0, pen_abs = 0;
            if (j.prior) {
                prior = B.alloc<int32_t>(w1 * w1 * w1 + gb::kPriorPad);
                h2d(prior, j.prior, w1 * w1 * w1, s);
                launch_fill_u32(reinterpret_cast<uint32_t*>(prior + w1 * w1 * w1), gb::kPriorPad,
                                0u, s);
                for (size_t e = 0; e < w1 * w1 * w1; ++e)
                    prior_abs = std::max(prior_abs, iabs64(j.prior[e]));
            }
            for (size_t e = 0; e < size_t(A) * A; ++e) pen_abs = std::max(pen_abs, iabs64(j.pen[e]));
            PackedRefs pr{};
            pr.nchunks = uint32_t((j.n_refs + kChunk - 1) / kChunk);
            pr.npad = pr.nchunks * kChunk;
            pr.nwords = words_for(j.r_max_len);
            uint32_t* words = B.alloc<uint32_t>(size_t(pr.nwords + 1) * pr.npad);
            uint32_t* len = B.alloc<uint32_t>(pr.npad);
            uint32_t* orig = B.alloc<uint32_t>(pr.npad);
            pr.words = words;
            pr.len = len;
            pr.orig = orig;
            B.packs.push_back({j.r_codes, j.r_offsets, j.n_refs, pr.nwords, pr.npad, words, len,
                               orig, 4u, j.r_max_len, TwinJob{}});
            BatchJob d{};
            d.q_codes = j.q_codes;
            d.q_off = j.q_offsets;
            d.n_queries = j.n_queries;
            d.n_refs = j.n_refs;
            d.refs = pr;
            d.pen = pen;
            d.A = A;
            d.gap_open = j.gap_open;
            d.gap_extend = j.gap_extend;
            d.prior = prior;
            d.w1 = uint32_t(w1);
            d.out = j.out;
            const int64_t Lmax = std::max(j.q_max_len, j.r_max_len);
            d.exact = 2 * Lmax * pen_abs + prior_abs + int64_t(j.gap_open) +
                          Lmax * int64_t(j.gap_extend) >= int64_t(INT32_MAX);
            {  // chunk groups that stay in L2 (see chunk_group), balanced
                const uint64_t chunk_bytes = uint64_t(kChunk) * (pr.nwords + 1) * 4;
                const uint64_t g0 = std::min<uint64_t>(
                    pr.nchunks, std::max<uint64_t>(1, (uint64_t(512) << 10) / chunk_bytes));
                const uint64_t ng = (pr.nchunks + g0 - 1) / g0;
                d.group = uint32_t((pr.nchunks + ng - 1) / ng);
            }
            d.profile = size_t(j.q_max_len) * gb::kRowBytes <= 48 * 1024;
            lds = std::max(lds, size_t(d.profile ? j.q_max_len : A) * gb::kRowBytes);
            int32_t big = 0;
            d.short_path = short_launch(j.q_max_len, j.r_max_len, A, d.exact, pen_abs, prior_abs,
                                        pr, &big);
            if (d.short_path) {
                d.big = big;
                d.prof_rows = j.q_max_len + gb::kFastMaxL;
                d.slice_rows = j.r_max_len + 1;
                const bool twin = twin_launch(j.n_queries, j.q_max_len, pen_abs, prior_abs, &big);
                lds = std::max(lds, twin ? size_t(d.prof_rows) * gb::kTwinRowBytes +
                                               size_t(d.slice_rows) * gb::kTwinSliceBytes
                                         : size_t(d.prof_rows) * gb::kRowBytes +
                                               size_t(d.slice_rows) * gb::kSliceStride *
                                                   sizeof(int32_t));
                if (twin) {
                    d.twin = 1;
                    d.big = big;
                    uint2* twins = B.alloc<uint2>(twin_capacity(j.n_queries));
                    d.twins = twins;
                    d.twin_sorted = B.alloc<uint32_t>(j.n_queries);
                    B.packs.back().tw = {j.q_offsets, j.n_queries, d.twin_sorted, twins};
                    B.any_twin = true;
                }
            }
            B.dev_jobs.push_back(d);
        }
        B.all_twin = !B.dev_jobs.empty() &&
                     std::all_of(B.dev_jobs.begin(), B.dev_jobs.end(),
                                 [](const BatchJob& d) { return d.twin != 0; });
        B.stage_ints = B.any_twin ? 2 * kChunk : kChunk;
        B.lds_bytes = size_t(B.stage_ints) * sizeof(int32_t) + lds;
        B.d_jobs = B.alloc<BatchJob>(std::max<size_t>(B.dev_jobs.size(), 1));
        B.d_next = B.alloc<unsigned long long>(1);

        // tile shape: memoised per workload signature, else tuned or the default
        const std::string sig = batch_signature(B.jobs);
        bool known = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(sig);
            if (it != g_tuned.end()) {
                B.tiles = it->second;
                known = true;
            }
        }
        if (!known) {
            B.tiles = {4, 256, uint32_t(B.cus) * 8};
            if (autotune && !B.dev_jobs.empty()) {
                hipEvent_t e0, e1;
                check(hipEventCreate(&e0), "event");
                check(hipEventCreate(&e1), "event");
                float best = 1e30f;
                BatchTiles best_t = B.tiles;
                for (uint32_t qpu : kTuneQpu)
                    for (uint32_t th : kTuneThreads)
                        for (uint32_t wpc : kTuneWgPerCu) {
                            B.tiles = {qpu, th, uint32_t(B.cus) * wpc};
                            B.layout(qpu, s);
                            B.enqueue(s);  // warm the candidate once
                            check(hipEventRecord(e0, s), "event");
                            B.enqueue(s);
                            check(hipEventRecord(e1, s), "event");
                            check(hipEventSynchronize(e1), "event");
                            float ms = 0;
                            check(hipEventElapsedTime(&ms, e0, e1), "event");
                            if (ms < best) {
                                best = ms;
                                best_t = B.tiles;
                            }
                        }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                B.tiles = best_t;
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tuned[sig] = B.tiles;
            }
        }
        B.layout(B.tiles.queries_per_unit, s);
        sync(s);

        // capture the batch once; every run replays the graph
        hipStream_t cs = nullptr;
        check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream");
        check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
        B.enqueue(cs);
        const hipError_t ce = hipStreamEndCapture(cs, &B.graph);
        (void)hipStreamDestroy(cs);
        check(ce, "end capture");
        check(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0), "graph instantiate");
        *plan = P.release();
        return ok(status);
    });
}

int32_t st_gapblock_batch_plan_run(st_batch_plan* plan, void* stream, st_status* status) {
    return guarded(status, [&] {
        if (!plan) throw HipError{ST_INVALID_ARGUMENT, "seqtree_hip: null batch plan"};
        LaunchTimer t("gapblock_batch_graph", static_cast<hipStream_t>(stream));
        check(hipGraphLaunch(plan->impl.exec, static_cast<hipStream_t>(stream)), "graph launch");
        return ok(status);
    });
}

int32_t st_gapblock_batch_plan_tiles(const st_batch_plan* plan, uint32_t* queries_per_unit,
                                     uint32_t* threads, uint32_t* workgroups,
                                     uint64_t* work_units) {
    if (!plan) return ST_INVALID_ARGUMENT;
    *queries_per_unit = plan->impl.tiles.queries_per_unit;
    *threads = plan->impl.tiles.threads;
    *workgroups = plan->impl.tiles.workgroups;
    *work_units = plan->impl.total_units;
    return ST_OK;
}

int32_t st_gapblock_batch_plan_twin_jobs(const st_batch_plan* plan, uint32_t* twin_jobs) {
    if (!plan || !twin_jobs) return ST_INVALID_ARGUMENT;
    *twin_jobs = uint32_t(std::count_if(plan->impl.dev_jobs.begin(), plan->impl.dev_jobs.end(),
                                        [](const BatchJob& d) { return d.twin != 0; }));
    return ST_OK;
}

void st_gapblock_batch_plan_destroy(st_batch_plan* plan) { delete plan; }

int32_t st_align_matrix_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                               uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                               const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                               const int32_t* sim, uint32_t alphabet_size, int32_t local,
                               int32_t dist, int32_t gap_open, int32_t gap_extend, int32_t* out,
                               void* stream, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    return guarded(status, [&] {
        require_device();
        bool overflow = false;
        gotoh_device(q_codes, q_offsets, n_queries, q_max_len, r_codes, r_offsets, n_refs,
                     r_max_len, sim, alphabet_size, local, dist, gap_open, gap_extend, out,
                     static_cast<hipStream_t>(stream), &overflow);
        if (overflow) {
            set_status(status, ST_OVERFLOW, "%s", overflow_message(dist != 0));
            return int(ST_OVERFLOW);
        }
        return ok(status);
    });
}

int32_t st_align_matrix(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                        const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                        const int32_t* sim, uint32_t alphabet_size, int32_t local, int32_t dist,
                        int32_t gap_open, int32_t gap_extend, int32_t* out, st_status* status) {
    if (alphabet_size == 0 || alphabet_size > kMaxAlphabet) return bad_alphabet(status, alphabet_size);
    if (n_queries == 0 || n_refs == 0) return ok(status);
    return guarded(status, [&] {
        require_device();
        if (n_queries == 1 && n_refs == 1 && !dist) {  // align_score: one pair
            SingleOut res;
            if (single_gotoh(q_codes + q_offsets[0], q_offsets[1] - q_offsets[0],
                             r_codes + r_offsets[0], r_offsets[1] - r_offsets[0], sim,
                             alphabet_size, local != 0, gap_open, gap_extend, false,
                             local ? "single_gotoh_local" : "single_gotoh_global", &res)) {
                out[0] = int32_t(res.score);  // |score| < 2^28 under the kernel's range bound
                return ok(status);
            }
        }
        const HostSet qh = host_set(q_codes, q_offsets, n_queries);
        const HostSet rh = host_set(r_codes, r_offsets, n_refs);
        HostCall call;
        const hipStream_t s = call.stream();
        Arena ar(s);
        ar.prefer(call.scratch(), kContextScratch);
        const DeviceSet qd = upload(ar, qh, call);
        const DeviceSet rd = upload(ar, rh, call);
        bool any_overflow = false;
        run_rows(qh, n_refs, out, call, ar, qd,
                 [&](const uint64_t* qoff, uint64_t, uint64_t rows, int32_t* dout) {
                     bool overflow = false;
                     gotoh_device(qd.bytes, qoff, rows, qh.max_len, rd.bytes, rd.off, n_refs,
                                  rh.max_len, sim, alphabet_size, local, dist, gap_open,
                                  gap_extend, dout, s, &overflow);
                     any_overflow = any_overflow || overflow;
                 });
        if (any_overflow) {
            set_status(status, ST_OVERFLOW, "%s", overflow_message(dist != 0));
            return int(ST_OVERFLOW);
        }
        return ok(status);
    });
}

int32_t st_hamming_matrix_device(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                                 uint32_t a_max_len, const uint8_t* b, const uint64_t* b_offsets,
                                 uint64_t n_b, uint32_t b_max_len, int32_t* out, void* stream,
                                 st_status* status) {
    return guarded(status, [&] {
        require_device();
        if (n_a == 0 || n_b == 0) return ok(status);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        Arena ar(s);
        uint64_t* res = ar.get<uint64_t>(2);
        launch_fill_u32(reinterpret_cast<uint32_t*>(res), 4, 0xFFFFFFFFu, s);
        launch_hamming_check(a_offsets, n_a, b_offsets, n_b, res, s);
        launched("hamming_check");
        uint64_t h[2];
        d2h(h, res, 2, s);
        sync(s);
        if (h[0] != ~0ull || h[1] != ~0ull) {
            const uint64_t i = h[0] != ~0ull ? 0 : h[1];
            const uint64_t k = h[0] != ~0ull ? h[0] : 0;
            uint64_t ao[2], bo[2];
            d2h(ao, a_offsets + i, 2, s);
            d2h(bo, b_offsets + k, 2, s);
            sync(s);
            return hamming_mismatch(status, ao[1] - ao[0], bo[1] - bo[0]);
        }
        hamming_device(a, a_offsets, n_a, a_max_len, b, b_offsets, n_b, b_max_len, out, s);
        return ok(status);
    });
}

int32_t st_hamming_matrix(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                          const uint8_t* b, const uint64_t* b_offsets, uint64_t n_b, int32_t* out,
                          st_status* status) {
    const HostSet ah = host_set(a, a_offsets, n_a);
    const HostSet bh = host_set(b, b_offsets, n_b);
    uint64_t i = 0, k = 0;
    if (first_mismatch(ah, bh, &i, &k))
        return hamming_mismatch(status, a_offsets[i + 1] - a_offsets[i],
                                b_offsets[k + 1] - b_offsets[k]);
    if (n_a == 0 || n_b == 0) return ok(status);
    return guarded(status, [&] {
        require_device();
        SingleArgs sa;
        if (n_a == 1 && n_b == 1 &&
            single_args(sa, a + a_offsets[0], a_offsets[1] - a_offsets[0], b + b_offsets[0],
                        b_offsets[1] - b_offsets[0], kSingleMaxR, 0, true)) {  // hamming
            SingleCall one;
            out[0] = int32_t(one.run(SingleKind::Hamming, sa, "single_hamming").score);
            return ok(status);
        }
        HostCall call;
        const hipStream_t s = call.stream();
        Arena ar(s);
        ar.prefer(call.scratch(), kContextScratch);
        const DeviceSet ad = upload(ar, ah, call);
        const DeviceSet bd = upload(ar, bh, call);
        run_rows(ah, n_b, out, call, ar, ad,
                 [&](const uint64_t* aoff, uint64_t, uint64_t rows, int32_t* dout) {
                     hamming_device(ad.bytes, aoff, rows, ah.max_len, bd.bytes, bd.off, n_b,
                                    bh.max_len, dout, s);
                 });
        return ok(status);
    });
}

int32_t st_levenshtein_matrix_device(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                                     uint32_t a_max_len, const uint8_t* b,
                                     const uint64_t* b_offsets, uint64_t n_b, uint32_t b_max_len,
                                     int32_t* out, void* stream, st_status* status) {
    return guarded(status, [&] {
        require_device();
        levenshtein_device(a, a_offsets, n_a, a_max_len, b, b_offsets, n_b, b_max_len, out,
                           static_cast<hipStream_t>(stream));
        return ok(status);
    });
}

int32_t st_levenshtein_matrix(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                              const uint8_t* b, const uint64_t* b_offsets, uint64_t n_b,
                              int32_t* out, st_status* status) {
    if (n_a == 0 || n_b == 0) return ok(status);
    return guarded(status, [&] {
        require_device();
        if (n_a == 1 && n_b == 1) {  // levenshtein: rows over the shorter string
            const uint8_t *x = a + a_offsets[0], *y = b + b_offsets[0];
            uint64_t lx = a_offsets[1] - a_offsets[0], ly = b_offsets[1] - b_offsets[0];
            if (ly < lx) {
                std::swap(x, y);
                std::swap(lx, ly);
            }
            SingleArgs sa;
            if (single_args(sa, x, lx, y, ly, kSingleMaxQ, 0)) {
                SingleCall one;
                out[0] = int32_t(one.run(SingleKind::Levenshtein, sa, "single_levenshtein").score);
                return ok(status);
            }
        }
        const HostSet ah = host_set(a, a_offsets, n_a);
        const HostSet bh = host_set(b, b_offsets, n_b);
        HostCall call;
        const hipStream_t s = call.stream();
        Arena ar(s);
        ar.prefer(call.scratch(), kContextScratch);
        const DeviceSet ad = upload(ar, ah, call);
        const DeviceSet bd = upload(ar, bh, call);
        run_rows(ah, n_b, out, call, ar, ad,
                 [&](const uint64_t* aoff, uint64_t, uint64_t rows, int32_t* dout) {
                     levenshtein_device(ad.bytes, aoff, rows, ah.max_len, bd.bytes, bd.off, n_b,
                                        bh.max_len, dout, s);
                 });
        return ok(status);
    });
}

}  // extern "C"
