// Per-call preparation kernels: reference packing (chunk-local length sort + transposed
// word layout) and query bucketing by length.

#include "st_internal.hpp"

namespace st {
namespace {

constexpr uint32_t kPackThreads = 512;  // kChunk / 2: one compare-exchange per thread per stage

// One workgroup per chunk of kChunk references. Sorts the chunk's slots by (length, index)
// with a bitonic network in LDS, then writes len/orig and the residues of each slot as
// little-endian words, word-major, so scoring waves read 64 consecutive slots per load.
__global__ __launch_bounds__(kPackThreads) void pack_refs_kernel(
    const uint8_t* __restrict__ bytes, const uint64_t* __restrict__ off, uint64_t n,
    uint32_t nwords, uint32_t npad, uint32_t* __restrict__ words, uint32_t* __restrict__ len_out,
    uint32_t* __restrict__ orig_out, uint32_t scale) {
    __shared__ uint64_t keys[kChunk];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = uint64_t(blockIdx.x) * kChunk;

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

    for (uint32_t s = tid; s < kChunk; s += kPackThreads) {
        const uint64_t key = keys[s];
        const uint32_t t = uint32_t(key & 2047u);
        const uint64_t k = base + t;
        const bool valid = k < n;
        const uint32_t L = valid ? uint32_t(key >> 11) : 0u;
        const uint64_t slot = base + s;
        len_out[slot] = L;
        orig_out[slot] = valid ? uint32_t(k) : kNoSlot;
        const uint8_t* src = bytes + (valid ? off[k] : 0);
        // bytes past the end: 0 for raw bytes, code 31 (byte offset 124) for scale 4
        const uint32_t pad = scale == 4 ? 124u : 0u;
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

// One workgroup: a histogram of the lengths in LDS, prefix sums of the counts and of the
// twins per length, a scatter of the query indices by length into `sorted`, then the twins.
// The order inside one length follows the LDS atomics, which changes which queries share a
// workgroup but no score. Lengths past kTwinMaxLen (which the host's bound excludes) are
// clamped so that no access leaves the tables.
constexpr uint32_t kTwinThreads = 1024;
constexpr uint32_t kTwinBins = kTwinMaxLen + 1;

__device__ void twin_list(const uint64_t* __restrict__ off, uint64_t n,
                          uint32_t* __restrict__ sorted, uint2* __restrict__ twins, uint64_t cap) {
    __shared__ uint32_t count[kTwinBins], cursor[kTwinBins], start[kTwinBins];
    __shared__ uint64_t tstart[kTwinBins + 1];
    const uint32_t tid = threadIdx.x;
    for (uint32_t b = tid; b < kTwinBins; b += kTwinThreads) {
        count[b] = 0;
        cursor[b] = 0;
    }
    __syncthreads();
    for (uint64_t i = tid; i < n; i += kTwinThreads) {
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
    for (uint64_t i = tid; i < n; i += kTwinThreads) {
        const uint64_t len64 = off[i + 1] - off[i];
        const uint32_t len = uint32_t(len64 < kTwinBins ? len64 : kTwinBins - 1);
        sorted[start[len] + atomicAdd(&cursor[len], 1u)] = uint32_t(i);
    }
    __syncthreads();  // `sorted` is complete and visible to the whole workgroup
    const uint64_t total = tstart[kTwinBins];
    for (uint64_t p = tid; p < cap; p += kTwinThreads) {
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

__global__ __launch_bounds__(kTwinThreads) void twin_queries_kernel(
    const uint64_t* __restrict__ off, uint64_t n, uint32_t* __restrict__ sorted,
    uint2* __restrict__ twins, uint64_t cap) {
    twin_list(off, n, sorted, twins, cap);
}

// Workgroup j: the twin list of batch job j (if it has twin rows).
__global__ __launch_bounds__(kTwinThreads) void twin_queries_batch_kernel(
    const BatchJob* __restrict__ jobs) {
    const BatchJob& J = jobs[blockIdx.x];
    if (!J.twin || J.n_queries == 0) return;
    twin_list(J.q_off, J.n_queries, J.twin_sorted, const_cast<uint2*>(J.twins),
              twin_capacity(J.n_queries));
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

void launch_pack_refs(const uint8_t* bytes, const uint64_t* off, uint64_t n, uint32_t nwords,
                      uint32_t npad, uint32_t* words, uint32_t* len, uint32_t* orig,
                      uint32_t scale, hipStream_t s) {
    const uint32_t chunks = npad / kChunk;
    if (chunks == 0) return;
    hipLaunchKernelGGL(pack_refs_kernel, dim3(chunks), dim3(kPackThreads), 0, s, bytes, off, n,
                       nwords, npad, words, len, orig, scale);
}

void launch_classify(const uint64_t* off, uint64_t n, const ClassBounds& b, uint32_t* counts,
                     uint32_t* lists, hipStream_t s) {
    if (n == 0) return;
    const uint32_t blocks = uint32_t((n + 255) / 256);
    hipLaunchKernelGGL(classify_kernel, dim3(blocks), dim3(256), 0, s, off, n, b, counts, lists);
}

void launch_twin_queries(const uint64_t* off, uint64_t n, uint32_t* sorted, uint2* twins,
                         hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(twin_queries_kernel, dim3(1), dim3(kTwinThreads), 0, s, off, n, sorted,
                       twins, twin_capacity(n));
}

void launch_twin_queries_batch(const BatchJob* jobs, uint32_t n_jobs, hipStream_t s) {
    if (n_jobs == 0) return;
    hipLaunchKernelGGL(twin_queries_batch_kernel, dim3(n_jobs), dim3(kTwinThreads), 0, s, jobs);
}

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    const uint32_t blocks = uint32_t((n + 255) / 256);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, n, v);
}

}  // namespace st
