// Per-call preparation kernels: reference packing (chunk-local length sort + transposed
// word layout) and query bucketing by length.

#include "st_internal.hpp"

namespace st {
namespace {

constexpr uint32_t kPackThreads = 512;  // kChunk / 2: one compare-exchange per thread per stage
constexpr uint32_t kSortBins = 257;     // counting sort: lengths 0..255, then the padding slots

// Twin pairing of a query set (see TwinJob in st_internal.hpp), with THREADS threads: an LDS
// histogram of the lengths, prefix sums of the counts and of the twins per length, a
// scatter of the query indices by length into `sorted`, then the twins. The order inside one
// length follows the LDS atomics, which changes which queries share a workgroup but no
// score. Lengths past kTwinMaxLen (which the host's bound excludes) are clamped so that no
// access leaves the tables.
constexpr uint32_t kTwinBins = kTwinMaxLen + 1;

template <uint32_t THREADS>
__device__ void twin_list(const uint64_t* __restrict__ off, uint64_t n,
                          uint32_t* __restrict__ sorted, uint2* __restrict__ twins, uint64_t cap) {
    __shared__ uint32_t count[kTwinBins], cursor[kTwinBins], start[kTwinBins];
    __shared__ uint64_t tstart[kTwinBins + 1];
    const uint32_t tid = threadIdx.x;
    for (uint32_t b = tid; b < kTwinBins; b += THREADS) {
        count[b] = 0;
        cursor[b] = 0;
    }
    __syncthreads();
    for (uint64_t i = tid; i < n; i += THREADS) {
        const uint64_t len = off[i + 1] - off[i];
        atomicAdd(&count[len < kTwinBins ? len : kTwinBins - 1], 1u);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t s = 0;
        uint64_t t = 0;
        for (uint32_t b = 0; b < kTwinBins; ++b) {
            start[b] = s;
            tstart[b] = t;
            s += count[b];
            t += (count[b] + 1) / 2;
        }
        tstart[kTwinBins] = t;
    }
    __syncthreads();
    for (uint64_t i = tid; i < n; i += THREADS) {
        const uint64_t len64 = off[i + 1] - off[i];
        const uint32_t len = uint32_t(len64 < kTwinBins ? len64 : kTwinBins - 1);
        sorted[start[len] + atomicAdd(&cursor[len], 1u)] = uint32_t(i);
    }
    __syncthreads();  // `sorted` is complete and visible to the whole workgroup
    const uint64_t total = tstart[kTwinBins];
    for (uint64_t p = tid; p < cap; p += THREADS) {
        uint2 v = make_uint2(kNoSlot, kNoSlot);
        if (p < total) {
            uint32_t b = 0;
            while (tstart[b + 1] <= p) ++b;  // the length whose twins include p
            const uint32_t j = uint32_t(p - tstart[b]);
            const uint32_t qa = sorted[start[b] + 2 * j];
            v = make_uint2(qa, 2 * j + 1 < count[b] ? sorted[start[b] + 2 * j + 1] : qa);
        }
        twins[p] = v;
    }
}

// One workgroup per chunk of kChunk references (plus, when tw.twins is set, one last
// workgroup that builds that query set's twin list, so the pairing costs no launch of its
// own). Orders the chunk's slots by length -- a counting sort in LDS when every length is
// below 256 (the order inside one length is then that of the LDS atomics, which no result
// depends on), else a bitonic network on (length, index) -- with the padding slots last,
// then writes len/orig and the residues of each slot as little-endian words, word-major, so
// scoring waves read 64 consecutive slots per load. A slot's words come from aligned dword
// loads through a bounds-checked buffer resource over the byte array (v_alignbyte realigns
// them), or byte by byte when the array is 2^31 bytes or more.
__global__ __launch_bounds__(kPackThreads) void pack_refs_kernel(
    const uint8_t* __restrict__ bytes, const uint64_t* __restrict__ off, uint64_t n,
    uint32_t nwords, uint32_t npad, uint32_t* __restrict__ words, uint32_t* __restrict__ len_out,
    uint32_t* __restrict__ orig_out, uint32_t scale, uint32_t max_len, TwinJob tw) {
    __shared__ uint64_t keys[kChunk];
    __shared__ uint32_t bins[kSortBins];
    __shared__ uint32_t seg_counts[(kChunk / 64) * kSortBins];
    const uint32_t tid = threadIdx.x;
    if (tw.twins != nullptr && blockIdx.x == gridDim.x - 1) {
        twin_list<kPackThreads>(tw.off, tw.n, tw.sorted, tw.twins, twin_capacity(tw.n));
        return;
    }
    const uint64_t base = uint64_t(blockIdx.x) * kChunk;

    if (max_len < kSortBins - 1) {
        // stable counting sort: bin = length (clamped), padding slots in the last bin. The
        // chunk is 16 segments of 64 slots (one wave each, twice over); a slot's position is
        // its bin's start + the same bin's count in the segments before + its rank among the
        // lower lanes of its segment with that bin -- the (length, index) order of the
        // comparison sort below.
        constexpr uint32_t kSegs = kChunk / 64;
        for (uint32_t e = tid; e < kSegs * kSortBins; e += kPackThreads) seg_counts[e] = 0;
        __syncthreads();
        const uint32_t lane = tid & 63u, wave = tid >> 6;
        uint32_t bin[kChunk / kPackThreads], rank[kChunk / kPackThreads];
#pragma unroll
        for (uint32_t u = 0; u < kChunk / kPackThreads; ++u) {
            const uint32_t seg = wave + u * (kPackThreads / 64);
            const uint64_t k = base + seg * 64 + lane;
            const uint64_t L = k < n ? (off[k + 1] - off[k]) : 0;
            bin[u] = k < n ? uint32_t(L < kSortBins - 1 ? L : kSortBins - 2) : kSortBins - 1;
            rank[u] = 0;
            uint64_t todo = ~0ull;
            while (todo) {  // once per distinct bin among the segment's 64 slots
                const uint32_t leader = uint32_t(__builtin_ctzll(todo));
                const uint32_t b = uint32_t(__builtin_amdgcn_readlane(int(bin[u]), int(leader)));
                const uint64_t same = __ballot(bin[u] == b);
                if (bin[u] == b) rank[u] = uint32_t(__popcll(same & ((1ull << lane) - 1)));
                if (lane == leader) seg_counts[seg * kSortBins + b] = uint32_t(__popcll(same));
                todo &= ~same;
            }
        }
        __syncthreads();
        for (uint32_t b = tid; b < kSortBins; b += kPackThreads) {  // prefix over segments
            uint32_t acc = 0;
            for (uint32_t g = 0; g < kSegs; ++g) {
                const uint32_t c = seg_counts[g * kSortBins + b];
                seg_counts[g * kSortBins + b] = acc;
                acc += c;
            }
            bins[b] = acc;
        }
        __syncthreads();
        {  // bin starts: exclusive prefix of the bins (Hillis-Steele, kSortBins <= threads)
            static_assert(kSortBins <= kPackThreads, "one thread per bin");
            const uint32_t own = tid < kSortBins ? bins[tid] : 0u;
            uint32_t x = own;
            for (uint32_t d = 1; d < kSortBins; d <<= 1) {
                const uint32_t y = (tid < kSortBins && tid >= d) ? bins[tid - d] : 0u;
                __syncthreads();
                x += y;
                if (tid < kSortBins) bins[tid] = x;
                __syncthreads();
            }
            if (tid < kSortBins) bins[tid] = x - own;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kChunk / kPackThreads; ++u) {
            const uint32_t seg = wave + u * (kPackThreads / 64);
            const uint32_t t = seg * 64 + lane;
            const uint64_t k = base + t;
            const uint64_t L = k < n ? (off[k + 1] - off[k]) : 0xFFFFFFFFull;
            keys[bins[bin[u]] + seg_counts[seg * kSortBins + bin[u]] + rank[u]] = (L << 11) | t;
        }
        __syncthreads();
    } else {
        for (uint32_t t = tid; t < kChunk; t += kPackThreads) {
            const uint64_t k = base + t;
            // padding slots sort last: their key length is 2^32 - 1, above any real length
            const uint64_t L = k < n ? (off[k + 1] - off[k]) : 0xFFFFFFFFull;
            keys[t] = (L << 11) | t;
        }
        __syncthreads();
        for (uint32_t size = 2; size <= kChunk; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                const uint32_t t = tid;
                const uint32_t i = 2 * stride * (t / stride) + (t % stride);
                const uint32_t j = i + stride;
                const bool ascending = (i & size) == 0;
                const uint64_t a = keys[i], b = keys[j];
                if ((a > b) == ascending) {
                    keys[i] = b;
                    keys[j] = a;
                }
                __syncthreads();
            }
        }
    }

    // the byte array through a buffer resource from a 4-byte-aligned base (reads past its
    // end return zeros)
    const uint64_t total = off[n];
    const uint32_t lead = uint32_t(reinterpret_cast<uintptr_t>(bytes) & 3u);
    const bool dwords = total + lead + 4 < (uint64_t(1) << 31);
    // the range ends on the dword holding the last byte (an aligned dword never crosses a
    // page), so a range check on whole dwords still returns every real byte
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(bytes - lead), 0, dwords ? int((total + lead + 3) & ~uint64_t(3)) : 0,
        0x00020000);
    // bytes past the end: 0 for raw bytes, code 31 (byte offset 124) for scale 4
    const uint32_t pad = scale == 4 ? 124u : 0u;
    const uint32_t pad4 = pad * 0x01010101u;
    for (uint32_t s = tid; s < kChunk; s += kPackThreads) {
        const uint64_t key = keys[s];
        const uint32_t t = uint32_t(key & 2047u);
        const uint64_t k = base + t;
        const bool valid = k < n;
        const uint32_t L = valid ? uint32_t(key >> 11) : 0u;
        const uint64_t slot = base + s;
        len_out[slot] = L;
        orig_out[slot] = valid ? uint32_t(k) : kNoSlot;
        const uint64_t o = valid ? off[k] : 0;
        if (dwords) {
            const uint32_t ob = uint32_t(o) + lead;  // byte offset from the aligned base
            const uint32_t sh = ob & 3u;
            uint32_t lo = uint32_t(__builtin_amdgcn_raw_buffer_load_b32(rsrc, ob & ~3u, 0, 0));
            for (uint32_t w = 0; w <= nwords; ++w) {
                const uint32_t hi =
                    uint32_t(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (ob & ~3u) + 4 * w + 4, 0, 0));
                uint32_t word = __builtin_amdgcn_alignbyte(hi, lo, sh);
                if (scale == 4) word = (word << 2) & 0xFCFCFCFCu;
                const uint32_t have = L > 4 * w ? L - 4 * w : 0u;  // real residues in this word
                const uint32_t keep = have >= 4 ? 0xFFFFFFFFu : (1u << (8 * have)) - 1u;
                words[uint64_t(w) * npad + slot] = (word & keep) | (pad4 & ~keep);
                lo = hi;
            }
        } else {
            const uint8_t* src = bytes + o;
            for (uint32_t w = 0; w <= nwords; ++w) {
                uint32_t word = 0;
                const uint32_t p = 4 * w;
#pragma unroll
                for (uint32_t b = 0; b < 4; ++b)
                    word |= (p + b < L ? uint32_t(src[p + b]) * scale : pad) << (8 * b);
                words[uint64_t(w) * npad + slot] = word;
            }
        }
    }
}

__global__ void classify_kernel(const uint64_t* __restrict__ off, uint64_t n, ClassBounds b,
                                uint32_t* __restrict__ counts, uint32_t* __restrict__ lists) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t m = off[i + 1] - off[i];
    int c = 0;
    while (c < b.count - 1 && m > b.bound[c]) ++c;
    const uint32_t pos = atomicAdd(&counts[c], 1u);
    lists[uint64_t(c) * n + pos] = uint32_t(i);
}

// Per-row histograms of an int32 matrix block: workgroup (row, segment) counts its segment's
// values in LDS -- bin 0 below zero, bin 1 + v for 0 <= v < H, bin H + 1 at H or above -- and
// adds the non-zero bins to the row's global histogram.
constexpr uint32_t kHistThreads = 256;
constexpr uint32_t kHistSegment = 1u << 16;

__global__ __launch_bounds__(kHistThreads) void score_histogram_kernel(
    const int32_t* __restrict__ m, uint64_t cols, uint32_t H, unsigned long long* __restrict__ out) {
    extern __shared__ uint32_t bins[];
    const uint32_t tid = threadIdx.x;
    const uint64_t row = blockIdx.y;
    const uint64_t c0 = uint64_t(blockIdx.x) * kHistSegment;
    const uint64_t c1 = c0 + kHistSegment < cols ? c0 + kHistSegment : cols;
    for (uint32_t b = tid; b < H + 2; b += kHistThreads) bins[b] = 0;
    __syncthreads();
    const int32_t* v = m + row * cols;
    // scores outside [0, H) are counted in registers: at a low ceiling nearly every score is
    // above it, and LDS atomics on one address serialise across the wave
    uint32_t below = 0, above = 0;
    for (uint64_t c = c0 + tid; c < c1; c += kHistThreads) {
        const int32_t x = v[c];
        if (x < 0) ++below;
        else if (uint32_t(x) >= H) ++above;
        else atomicAdd(&bins[uint32_t(x) + 1], 1u);
    }
    if (below) atomicAdd(&bins[0], below);
    if (above) atomicAdd(&bins[H + 1], above);
    __syncthreads();
    unsigned long long* o = out + row * (H + 2);
    for (uint32_t b = tid; b < H + 2; b += kHistThreads)
        if (bins[b]) atomicAdd(&o[b], (unsigned long long)bins[b]);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void launch_pack_refs(const uint8_t* bytes, const uint64_t* off, uint64_t n, uint32_t nwords,
                      uint32_t npad, uint32_t* words, uint32_t* len, uint32_t* orig,
                      uint32_t scale, uint32_t max_len, const TwinJob& tw, hipStream_t s) {
    const uint32_t chunks = npad / kChunk;
    if (chunks == 0 && tw.twins == nullptr) return;
    hipLaunchKernelGGL(pack_refs_kernel, dim3(chunks + (tw.twins ? 1 : 0)), dim3(kPackThreads), 0,
                       s, bytes, off, n, nwords, npad, words, len, orig, scale, max_len, tw);
}

void launch_classify(const uint64_t* off, uint64_t n, const ClassBounds& b, uint32_t* counts,
                     uint32_t* lists, hipStream_t s) {
    if (n == 0) return;
    const uint32_t blocks = uint32_t((n + 255) / 256);
    hipLaunchKernelGGL(classify_kernel, dim3(blocks), dim3(256), 0, s, off, n, b, counts, lists);
}

void launch_score_histogram(const int32_t* m, uint64_t rows, uint64_t cols, uint32_t H,
                            unsigned long long* out, hipStream_t s) {
    if (rows == 0 || cols == 0) return;
    const dim3 grid(uint32_t((cols + kHistSegment - 1) / kHistSegment), uint32_t(rows));
    hipLaunchKernelGGL(score_histogram_kernel, grid, dim3(kHistThreads),
                       size_t(H + 2) * sizeof(uint32_t), s, m, cols, H, out);
}

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    const uint32_t blocks = uint32_t((n + 255) / 256);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
}

}  // namespace st
