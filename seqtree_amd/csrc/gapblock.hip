// Single-gap-block scoring kernel (reference: src/gapblock.cpp:39-62 `cell`, driven by
// src/gapblock.cpp:66-133).
//
// Workgroup = (query i, reference chunk c). The query's penalty profile (row t holds
// pen[q[t]][b] and its negation, see gapblock_core.hpp) sits in LDS -- or, for very long
// queries, the A x A penalty table in the same row format -- so every lookup is one ds_read
// whose row is wave-uniform. Each lane owns one reference slot; its residues stream in as
// 4-byte words of byte offsets (coalesced across the wave), and the "reference longer"
// diagonal is realigned in registers with v_alignbyte instead of a second byte gather.
//
// With L = min(m, n), d = |m - n|, p0(t) the diagonal-0 penalty and pd(t) the diagonal-d
// penalty at block-start position t, the reference computes
//     best = min_{i in [0, L]}  P0[i] + (Sd - Pd[i]) + prior[i],   Sd = Pd[L]
// (prefix P0, suffix Sd - Pd[i]). Without int32 overflow that equals Sd + min_i (P0[i] - Pd[i]
// + prior[i]), which one forward pass computes (2 lookups per position). When an int32 sum
// can wrap, the exact variant evaluates (P0[i] + (Sd - Pd[i])) + prior[i] in the reference's
// order with wrapping arithmetic after a first pass for Sd, so the chosen minimum is the one
// the reference's int32 code produces.

#include "gapblock_core.hpp"

namespace st {
namespace {

using gb::CsrWords;
using gb::Look;
using gb::PackedWords;
using gb::pair_score;
using gb::pair_score_fast;

// Appends (query, ref, score) for every lane with hit set: one atomic per wave.
__device__ __forceinline__ void append_hit(bool hit, uint32_t qi, uint32_t k, int32_t score,
                                           const GapblockHits& h) {
    const unsigned long long mask = __ballot(hit);
    if (mask == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll(mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(h.count, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    if (hit) {
        const unsigned long long idx = base + __popcll(mask & ((1ull << lane) - 1));
        if (idx < h.capacity) {
            h.query[idx] = qi;
            h.ref[idx] = k;
            h.score[idx] = score;
        }
    }
}

// Histogram output: the workgroup's LDS bins (kHistFusedBins per query row) count the scores
// in [0, H); scores below zero and at or above H are counted in each lane's registers (at a
// low ceiling nearly every score is above it, and LDS atomics on one address serialise across
// the wave) and added once at the end, when the non-zero bins go to the global row.
__device__ __forceinline__ void hist_add(uint32_t* bins, bool valid, int32_t score, uint32_t H,
                                         uint32_t& below, uint32_t& above) {
    if (!valid) return;
    if (score < 0) ++below;
    else if (uint32_t(score) >= H) ++above;
    else atomicAdd(&bins[uint32_t(score) + 1], 1u);
}

// After a barrier that follows every hist_add and the register counts' atomics.
__device__ __forceinline__ void hist_flush(const uint32_t* bins, uint32_t H,
                                           unsigned long long* row, uint32_t tid) {
    for (uint32_t b = tid; b < H + 2; b += kThreads)
        if (bins[b]) atomicAdd(&row[b], (unsigned long long)bins[b]);
}

__device__ __forceinline__ void hist_tail(uint32_t* bins, uint32_t H, uint32_t below,
                                          uint32_t above) {
    if (below) atomicAdd(&bins[0], below);
    if (above) atomicAdd(&bins[H + 1], above);
}

// OUT: kOutDense writes the N x K matrix, kOutHits appends hits under the cutoff, kOutHist
// adds each query row's scores to its histogram.
constexpr int kOutDense = 0, kOutHits = 1, kOutHist = 2;

template <int OUT>
constexpr uint32_t stage_ints(uint32_t rows) {
    return OUT == kOutDense ? rows * kChunk : OUT == kOutHist ? rows * kHistFusedBins : 1u;
}

// SHORT: the short-launch layout (requires PROFILE && !EXACT): profile rows, then the
// prior slice, both in the dynamic LDS.
// A workgroup owns query i and chunks [g * group, min((g + 1) * group, nchunks)): the
// query's tables are built once and reused across the group's chunks. Workgroups are
// numbered group-major (all queries of group 0 first), so the workgroups resident at one
// time stream the same few reference chunks and find them in L2.
template <bool PROFILE, bool EXACT, int OUT, bool SHORT>
__global__ __launch_bounds__(kThreads) void gapblock_kernel(GapblockArgs a, uint64_t q_base,
                                                            uint32_t nq) {
    extern __shared__ int32_t lds[];
    __shared__ int32_t stage[stage_ints<OUT>(1)];

    const uint32_t tid = threadIdx.x;
    const uint32_t nch = a.refs.nchunks;
    const uint32_t g = blockIdx.x / nq;
    const uint64_t i = q_base + blockIdx.x % nq;
    const uint32_t c_end = (g + 1) * a.group < nch ? (g + 1) * a.group : nch;
    const uint64_t q0 = a.q_off[i];
    const uint32_t m = uint32_t(a.q_off[i + 1] - q0);
    const uint8_t* q = a.q_codes + q0;
    const uint32_t A = a.A;
    int32_t* slice = lds + (SHORT ? a.prof_rows : 0u) * gb::kRowInts;

    if constexpr (SHORT) {
        gb::build_rows(lds, a.prof_rows, m, a.pen, A, [&](uint32_t t) { return uint32_t(q[t]); },
                       tid, kThreads);
        gb::build_slice(slice, a.slice_rows - 1, m, a.prior, a.w1, a.big, tid, kThreads);
    } else if constexpr (PROFILE) {
        gb::build_rows(lds, m, m, a.pen, A, [&](uint32_t t) { return uint32_t(q[t]); }, tid,
                       kThreads);
    } else {
        gb::build_rows(lds, A, A, a.pen, A, [](uint32_t r) { return r; }, tid, kThreads);
    }
    uint32_t* bins = reinterpret_cast<uint32_t*>(stage);
    uint32_t below = 0, above = 0;
    if constexpr (OUT == kOutHist)
        for (uint32_t b = tid; b < a.hist_bins + 2; b += kThreads) bins[b] = 0;
    __syncthreads();

    const Look<PROFILE> look{lds, q};
    // the next slot's column and length are loaded one pair ahead (across chunks too)
    uint32_t next_k = 0, next_n = 0;
    if (g * a.group < c_end) {
        next_k = a.refs.orig[g * a.group * kChunk + tid];
        next_n = a.refs.len[g * a.group * kChunk + tid];
    }
    for (uint32_t c = g * a.group; c < c_end; ++c) {
        for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
            const uint32_t slot = c * kChunk + it * kThreads + tid;
            const uint32_t k = next_k;
            const uint32_t n = next_n;
            {
                const uint32_t it2 = it + 1 < kSlotsPerThread ? it + 1 : 0;
                const uint32_t c2 = it + 1 < kSlotsPerThread ? c : c + 1;
                if (c2 < c_end) {
                    next_k = a.refs.orig[c2 * kChunk + it2 * kThreads + tid];
                    next_n = a.refs.len[c2 * kChunk + it2 * kThreads + tid];
                }
            }
            const PackedWords words{gb::global_ptr(a.refs.words) + slot, a.refs.npad};
            int32_t score = 0;
            if constexpr (SHORT) {
                score = gb::pair_score_short(
                    lds, slice, gb::buffer_words(a.refs.words, a.refs.npad, a.refs.nwords, slot),
                    k, m, n, a.prior, a.w1, a.gap_open, a.gap_extend);
            } else {
                // one reference length across the wave (padding lanes excepted): uniform path
                const uint32_t n0 = __builtin_amdgcn_readfirstlane(n);
                const bool uniform = __all(k == kNoSlot || n == n0);
                if (PROFILE && !EXACT && uniform && (m < n0 ? m : n0) <= gb::kFastMaxL) {
                    score = pair_score_fast(lds, words, m, n0, a.prior, a.w1, a.gap_open,
                                            a.gap_extend);
                } else if (k != kNoSlot) {
                    score = pair_score<PROFILE, EXACT>(look, words, m, n, a.prior, a.w1,
                                                       a.gap_open, a.gap_extend);
                }
            }
            if constexpr (OUT == kOutHits) {
                append_hit(k != kNoSlot && score <= a.hits.cutoff, uint32_t(i), k, score, a.hits);
            } else if constexpr (OUT == kOutHist) {
                hist_add(bins, k != kNoSlot, score, a.hist_bins, below, above);
            } else {
                if (k != kNoSlot) stage[k - c * kChunk] = score;
            }
        }
        if constexpr (OUT == kOutDense) {
            __syncthreads();
            const uint64_t col0 = uint64_t(c) * kChunk;
            const uint64_t lim64 = a.n_refs - col0;
            const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
            int32_t* row = a.out + i * a.n_refs + col0;
            for (uint32_t t = tid; t < lim; t += kThreads) row[t] = stage[t];
            __syncthreads();  // the stage is rewritten by the next chunk
        }
    }
    if constexpr (OUT == kOutHist) {
        hist_tail(bins, a.hist_bins, below, above);
        __syncthreads();
        hist_flush(bins, a.hist_bins, a.hist + i * (a.hist_bins + 2), tid);
    }
}

// Twin short launch: the short layout with two queries of one length per workgroup, packed
// into the 16-bit halves of the profile and the prior slice (twin rows, gapblock_core.hpp).
// Workgroup = (twin p, chunk group g), group-major like gapblock_kernel; twins[p] = (a, b)
// with b == a for the last query of a length with an odd count (one row written), and
// (kNoSlot, kNoSlot) past the last twin (the workgroup exits before any barrier).
template <int OUT>
__global__ __launch_bounds__(kThreads) void gapblock_twin_kernel(GapblockArgs a,
                                                                 const uint2* __restrict__ twins,
                                                                 uint32_t nt) {
    extern __shared__ int32_t lds[];
    __shared__ int32_t stage[stage_ints<OUT>(2)];

    const uint32_t tid = threadIdx.x;
    const uint32_t nch = a.refs.nchunks;
    const uint32_t g = blockIdx.x / nt;
    const uint2 tw = twins[blockIdx.x % nt];
    if (tw.x == kNoSlot) return;
    const uint32_t c_end = (g + 1) * a.group < nch ? (g + 1) * a.group : nch;
    const uint64_t ia = tw.x, ib = tw.y;
    const uint64_t qa0 = a.q_off[ia], qb0 = a.q_off[ib];
    const uint32_t m = uint32_t(a.q_off[ia + 1] - qa0);
    uint32_t* prof = reinterpret_cast<uint32_t*>(lds);
    uint32_t* slice = prof + a.prof_rows * gb::kTwinRowInts;
    gb::build_rows2(prof, a.prof_rows, m, a.pen, a.A, a.q_codes + qa0, a.q_codes + qb0, tid,
                    kThreads);
    gb::build_slice2(slice, a.slice_rows - 1, m, a.prior, a.w1, a.big, tid, kThreads);
    // histogram rows of a and b (b's unused when the twin is one query)
    uint32_t* bins_a = reinterpret_cast<uint32_t*>(stage);
    uint32_t* bins_b = bins_a + kHistFusedBins;
    uint32_t below_a = 0, above_a = 0, below_b = 0, above_b = 0;
    if constexpr (OUT == kOutHist)
        for (uint32_t b = tid; b < 2 * kHistFusedBins; b += kThreads) bins_a[b] = 0;
    __syncthreads();

    uint32_t next_k = 0, next_n = 0;
    if (g * a.group < c_end) {
        next_k = a.refs.orig[g * a.group * kChunk + tid];
        next_n = a.refs.len[g * a.group * kChunk + tid];
    }
    for (uint32_t c = g * a.group; c < c_end; ++c) {
        for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
            const uint32_t slot = c * kChunk + it * kThreads + tid;
            const uint32_t k = next_k;
            const uint32_t n = next_n;
            {
                const uint32_t it2 = it + 1 < kSlotsPerThread ? it + 1 : 0;
                const uint32_t c2 = it + 1 < kSlotsPerThread ? c : c + 1;
                if (c2 < c_end) {
                    next_k = a.refs.orig[c2 * kChunk + it2 * kThreads + tid];
                    next_n = a.refs.len[c2 * kChunk + it2 * kThreads + tid];
                }
            }
            int32_t sa = 0, sb = 0;
            gb::twin_score_short(prof, slice,
                                 gb::buffer_words(a.refs.words, a.refs.npad, a.refs.nwords, slot),
                                 k, m, n, a.gap_open, a.gap_extend, sa, sb);
            if constexpr (OUT == kOutHits) {
                append_hit(k != kNoSlot && sa <= a.hits.cutoff, uint32_t(ia), k, sa, a.hits);
                if (ib != ia)
                    append_hit(k != kNoSlot && sb <= a.hits.cutoff, uint32_t(ib), k, sb, a.hits);
            } else if constexpr (OUT == kOutHist) {
                hist_add(bins_a, k != kNoSlot, sa, a.hist_bins, below_a, above_a);
                hist_add(bins_b, k != kNoSlot && ib != ia, sb, a.hist_bins, below_b, above_b);
            } else if (k != kNoSlot) {
                stage[k - c * kChunk] = sa;
                stage[kChunk + k - c * kChunk] = sb;
            }
        }
        if constexpr (OUT == kOutDense) {
            __syncthreads();
            const uint64_t col0 = uint64_t(c) * kChunk;
            const uint64_t lim64 = a.n_refs - col0;
            const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
            int32_t* row_a = a.out + ia * a.n_refs + col0;
            for (uint32_t t = tid; t < lim; t += kThreads) row_a[t] = stage[t];
            if (ib != ia) {
                int32_t* row_b = a.out + ib * a.n_refs + col0;
                for (uint32_t t = tid; t < lim; t += kThreads) row_b[t] = stage[kChunk + t];
            }
            __syncthreads();  // the stage is rewritten by the next chunk
        }
    }
    if constexpr (OUT == kOutHist) {
        hist_tail(bins_a, a.hist_bins, below_a, above_a);
        hist_tail(bins_b, a.hist_bins, below_b, above_b);
        __syncthreads();
        hist_flush(bins_a, a.hist_bins, a.hist + ia * (a.hist_bins + 2), tid);
        if (ib != ia) hist_flush(bins_b, a.hist_bins, a.hist + ib * (a.hist_bins + 2), tid);
    }
}

// Pair-list mode: thread t scores (pair_query[t], pair_ref[t]) straight from the CSR sets with
// the A x A penalty table in LDS.
template <bool EXACT>
__global__ __launch_bounds__(kThreads) void gapblock_pairs_kernel(
    GapblockArgs a, const uint8_t* __restrict__ r_codes, const uint64_t* __restrict__ r_off,
    const uint32_t* __restrict__ pq, const uint32_t* __restrict__ pr, uint64_t n_pairs) {
    __shared__ int32_t pen[kMaxAlphabet * gb::kRowInts];
    const uint32_t A = a.A;
    gb::build_rows(pen, A, A, a.pen, A, [](uint32_t r) { return r; }, threadIdx.x, kThreads);
    __syncthreads();
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t t = uint64_t(blockIdx.x) * kThreads + threadIdx.x; t < n_pairs; t += stride) {
        const uint32_t qi = pq[t], k = pr[t];
        const uint64_t q0 = a.q_off[qi], r0 = r_off[k];
        const uint32_t m = uint32_t(a.q_off[qi + 1] - q0), n = uint32_t(r_off[k + 1] - r0);
        const Look<false> look{pen, a.q_codes + q0};
        a.out[t] = pair_score<false, EXACT>(look, CsrWords{r_codes + r0, n}, m, n, a.prior, a.w1,
                                            a.gap_open, a.gap_extend);
    }
}

template <bool PROFILE, bool EXACT, int OUT, bool SHORT>
void launch_variant(const GapblockArgs& a, size_t lds_bytes, hipStream_t s) {
    const uint64_t ngroups = (a.refs.nchunks + a.group - 1) / a.group;
    const uint64_t max_q = (uint64_t(1) << 30) / ngroups;  // launches under 2^30 workgroups
    for (uint64_t q0 = 0; q0 < a.n_queries; q0 += max_q) {
        const uint64_t nq = a.n_queries - q0 < max_q ? a.n_queries - q0 : max_q;
        hipLaunchKernelGGL((gapblock_kernel<PROFILE, EXACT, OUT, SHORT>),
                           dim3(uint32_t(nq * ngroups)), dim3(kThreads), lds_bytes, s, a, q0,
                           uint32_t(nq));
    }
}

template <int OUT>
void launch_mode(const GapblockArgs& a, bool use_profile, size_t lds_bytes, bool exact_wrap,
                 bool short_path, hipStream_t s) {
    if (short_path) {
        launch_variant<true, false, OUT, true>(a, lds_bytes, s);
    } else if (use_profile) {
        if (exact_wrap) launch_variant<true, true, OUT, false>(a, lds_bytes, s);
        else launch_variant<true, false, OUT, false>(a, lds_bytes, s);
    } else {
        if (exact_wrap) launch_variant<false, true, OUT, false>(a, lds_bytes, s);
        else launch_variant<false, false, OUT, false>(a, lds_bytes, s);
    }
}

}  // namespace

void launch_gapblock(const GapblockArgs& a, bool use_profile, uint32_t q_max_len, bool exact_wrap,
                     bool short_path, hipStream_t s) {
    if (a.n_queries == 0 || a.refs.nchunks == 0) return;
    // one table row (penalties and their negations) per query position or per residue; short
    // launches add zero rows and the prior slice
    const size_t lds =
        short_path ? size_t(a.prof_rows) * gb::kRowBytes +
                         size_t(a.slice_rows) * gb::kSliceStride * sizeof(int32_t)
                   : size_t(use_profile ? q_max_len : a.A) * gb::kRowBytes;
    const size_t lds_bytes = lds ? lds : sizeof(int32_t);
    if (a.hist) launch_mode<kOutHist>(a, use_profile, lds_bytes, exact_wrap, short_path, s);
    else if (a.hits.count) launch_mode<kOutHits>(a, use_profile, lds_bytes, exact_wrap, short_path, s);
    else launch_mode<kOutDense>(a, use_profile, lds_bytes, exact_wrap, short_path, s);
}

void launch_gapblock_twin(const GapblockArgs& a, const uint2* twins, uint64_t n_twins,
                          hipStream_t s) {
    if (n_twins == 0 || a.refs.nchunks == 0) return;
    const size_t lds = size_t(a.prof_rows) * gb::kTwinRowBytes +
                       size_t(a.slice_rows) * gb::kTwinSliceBytes;
    const uint64_t ngroups = (a.refs.nchunks + a.group - 1) / a.group;
    const uint64_t max_t = (uint64_t(1) << 30) / ngroups;  // launches under 2^30 workgroups
    for (uint64_t t0 = 0; t0 < n_twins; t0 += max_t) {
        const uint32_t nt = uint32_t(n_twins - t0 < max_t ? n_twins - t0 : max_t);
        const dim3 grid(uint32_t(nt * ngroups));
        if (a.hist)
            hipLaunchKernelGGL(gapblock_twin_kernel<kOutHist>, grid, dim3(kThreads), lds, s, a,
                               twins + t0, nt);
        else if (a.hits.count)
            hipLaunchKernelGGL(gapblock_twin_kernel<kOutHits>, grid, dim3(kThreads), lds, s, a,
                               twins + t0, nt);
        else
            hipLaunchKernelGGL(gapblock_twin_kernel<kOutDense>, grid, dim3(kThreads), lds, s, a,
                               twins + t0, nt);
    }
}

void launch_gapblock_pairs(const GapblockArgs& a, const uint8_t* r_codes, const uint64_t* r_off,
                           const uint32_t* pair_query, const uint32_t* pair_ref, uint64_t n_pairs,
                           bool exact_wrap, hipStream_t s) {
    if (n_pairs == 0) return;
    const uint64_t blocks64 = (n_pairs + kThreads - 1) / kThreads;
    const uint32_t blocks = uint32_t(blocks64 < 65536 ? blocks64 : 65536);
    if (exact_wrap)
        hipLaunchKernelGGL(gapblock_pairs_kernel<true>, dim3(blocks), dim3(kThreads), 0, s, a,
                           r_codes, r_off, pair_query, pair_ref, n_pairs);
    else
        hipLaunchKernelGGL(gapblock_pairs_kernel<false>, dim3(blocks), dim3(kThreads), 0, s, a,
                           r_codes, r_off, pair_query, pair_ref, n_pairs);
}

}  // namespace st
