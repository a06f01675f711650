// Single-gap-block scoring kernel (reference: src/gapblock.cpp:39-62 `cell`, driven by
// src/gapblock.cpp:66-133).
//
// Workgroup = (query i, reference chunk c). The query's penalty profile prof[t][b] =
// pen[q[t]][b] sits in LDS (or, for very long queries, the A x A penalty table), so every
// lookup is one ds_read whose row is wave-uniform. Each lane owns one reference slot; its
// residues stream in as 4-byte words (coalesced across the wave), and the "reference longer"
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

#include "st_internal.hpp"

#include <climits>

namespace st {
namespace {

template <bool PROFILE>
struct Look {
    const int32_t* lds;
    const uint8_t* q;
    uint32_t A;
    __device__ __forceinline__ int32_t operator()(uint32_t t, uint32_t rc) const {
        if constexpr (PROFILE) {
            return lds[t * A + rc];
        } else {
            return lds[uint32_t(q[t]) * A + rc];
        }
    }
};

// Reference residues as little-endian words: residues 4w..4w+3 of one sequence.
struct PackedWords {  // the chunk-packed layout, one slot
    const uint32_t* wp;
    uint32_t npad;
    __device__ __forceinline__ uint32_t operator()(uint32_t w) const {
        return wp[uint64_t(w) * npad];
    }
};

struct CsrWords {  // straight from the CSR bytes (pair-list mode)
    const uint8_t* r;
    uint32_t n;
    __device__ __forceinline__ uint32_t operator()(uint32_t w) const {
        uint32_t word = 0;
        const uint32_t p = 4 * w;
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b)
            if (p + b < n) word |= uint32_t(r[p + b]) << (8 * b);
        return word;
    }
};

// Visits block-start positions t = 0..L-1 in order, handing the two penalties to f.
template <bool PROFILE, class Words, class F>
__device__ __forceinline__ void walk(const Look<PROFILE>& look, const Words& words, uint32_t L,
                                     uint32_t d, bool q_longer, F&& f) {
    for (uint32_t w = 0; 4 * w < L; ++w) {
        const uint32_t x = words(w);
        uint32_t y = 0;
        if (!q_longer) {
            const uint32_t t0 = 4 * w + d;
            const uint32_t lo = words(t0 >> 2);
            const uint32_t hi = words((t0 >> 2) + 1);
            y = __builtin_amdgcn_alignbyte(hi, lo, t0 & 3u);
        }
#pragma unroll
        for (uint32_t kk = 0; kk < 4; ++kk) {
            const uint32_t t = 4 * w + kk;
            if (t < L) {
                const uint32_t rc = (x >> (8 * kk)) & 0xFFu;
                const int32_t p0 = look(t, rc);
                const int32_t pd =
                    q_longer ? look(t + d, rc) : look(t, (y >> (8 * kk)) & 0xFFu);
                f(t, p0, pd);
            }
        }
    }
}

// Gap-block score of one pair (query of length m behind `look`, reference of length n behind
// `words`). prior: the (w1)^3 cube or nullptr.
template <bool PROFILE, bool EXACT, class Words>
__device__ __forceinline__ int32_t pair_score(const Look<PROFILE>& look, const Words& words,
                                              uint32_t m, uint32_t n, const int32_t* prior,
                                              uint32_t w1, int32_t go, int32_t ge) {
    const uint32_t L = m < n ? m : n;
    const uint32_t d = m > n ? m - n : n - m;
    const bool q_longer = m >= n;
    const int32_t* prow =
        (prior && d) ? prior + (uint64_t(q_longer ? m : n) * w1 + d) * w1 : nullptr;
    if constexpr (!EXACT) {
        uint32_t P0 = 0, Pd = 0;
        int32_t best = prow ? prow[0] : 0;
        walk<PROFILE>(look, words, L, d, q_longer, [&](uint32_t t, int32_t p0, int32_t pd) {
            P0 += uint32_t(p0);
            Pd += uint32_t(pd);
            if (d) {
                const int32_t cand = int32_t(P0 - Pd + uint32_t(prow ? prow[t + 1] : 0));
                best = cand < best ? cand : best;
            }
        });
        return d == 0 ? int32_t(Pd)
                      : int32_t(Pd + uint32_t(best) + uint32_t(go) + (d - 1) * uint32_t(ge));
    } else {
        uint32_t Sd = 0;
        walk<PROFILE>(look, words, L, d, q_longer,
                      [&](uint32_t, int32_t, int32_t pd) { Sd += uint32_t(pd); });
        if (d == 0) return int32_t(Sd);
        uint32_t P0 = 0, Pd = 0;
        int32_t best = INT_MAX;
        // candidate i = t is formed before position t is consumed, i = L after the loop
        walk<PROFILE>(look, words, L, d, q_longer, [&](uint32_t t, int32_t p0, int32_t pd) {
            const uint32_t pr = prow ? uint32_t(prow[t]) : 0u;
            const int32_t cand = int32_t((P0 + (Sd - Pd)) + pr);
            best = cand < best ? cand : best;
            P0 += uint32_t(p0);
            Pd += uint32_t(pd);
        });
        const uint32_t pr = prow ? uint32_t(prow[L]) : 0u;
        const int32_t cand = int32_t((P0 + (Sd - Pd)) + pr);
        best = cand < best ? cand : best;
        return int32_t(uint32_t(best) + uint32_t(go) + (d - 1) * uint32_t(ge));
    }
}

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

template <bool PROFILE, bool EXACT, bool THRESH>
__global__ __launch_bounds__(kThreads) void gapblock_kernel(GapblockArgs a, uint64_t q_base) {
    extern __shared__ int32_t lds[];
    __shared__ int32_t stage[THRESH ? 1 : kChunk];

    const uint32_t tid = threadIdx.x;
    const uint32_t nch = a.refs.nchunks;
    const uint32_t c = blockIdx.x % nch;
    const uint64_t i = q_base + blockIdx.x / nch;
    const uint64_t q0 = a.q_off[i];
    const uint32_t m = uint32_t(a.q_off[i + 1] - q0);
    const uint8_t* q = a.q_codes + q0;
    const uint32_t A = a.A;

    if constexpr (PROFILE) {
        for (uint32_t e = tid; e < m * A; e += kThreads) {
            const uint32_t t = e / A;
            lds[e] = a.pen[uint32_t(q[t]) * A + (e - t * A)];
        }
    } else {
        for (uint32_t e = tid; e < A * A; e += kThreads) lds[e] = a.pen[e];
    }
    __syncthreads();

    const Look<PROFILE> look{lds, q, A};
    for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
        const uint32_t slot = c * kChunk + it * kThreads + tid;
        const uint32_t k = a.refs.orig[slot];
        int32_t score = 0;
        if (k != kNoSlot)
            score = pair_score<PROFILE, EXACT>(look, PackedWords{a.refs.words + slot, a.refs.npad},
                                               m, a.refs.len[slot], a.prior, a.w1, a.gap_open,
                                               a.gap_extend);
        if constexpr (THRESH) {
            append_hit(k != kNoSlot && score <= a.hits.cutoff, uint32_t(i), k, score, a.hits);
        } else {
            if (k != kNoSlot) stage[k - c * kChunk] = score;
        }
    }
    if constexpr (!THRESH) {
        __syncthreads();
        const uint64_t col0 = uint64_t(c) * kChunk;
        const uint64_t lim64 = a.n_refs - col0;
        const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
        int32_t* row = a.out + i * a.n_refs + col0;
        for (uint32_t t = tid; t < lim; t += kThreads) row[t] = stage[t];
    }
}

// Pair-list mode: thread t scores (pair_query[t], pair_ref[t]) straight from the CSR sets with
// the A x A penalty table in LDS.
template <bool EXACT>
__global__ __launch_bounds__(kThreads) void gapblock_pairs_kernel(
    GapblockArgs a, const uint8_t* __restrict__ r_codes, const uint64_t* __restrict__ r_off,
    const uint32_t* __restrict__ pq, const uint32_t* __restrict__ pr, uint64_t n_pairs) {
    __shared__ int32_t pen[kMaxAlphabet * kMaxAlphabet];
    const uint32_t A = a.A;
    for (uint32_t e = threadIdx.x; e < A * A; e += kThreads) pen[e] = a.pen[e];
    __syncthreads();
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t t = uint64_t(blockIdx.x) * kThreads + threadIdx.x; t < n_pairs; t += stride) {
        const uint32_t qi = pq[t], k = pr[t];
        const uint64_t q0 = a.q_off[qi], r0 = r_off[k];
        const uint32_t m = uint32_t(a.q_off[qi + 1] - q0), n = uint32_t(r_off[k + 1] - r0);
        const Look<false> look{pen, a.q_codes + q0, A};
        a.out[t] = pair_score<false, EXACT>(look, CsrWords{r_codes + r0, n}, m, n, a.prior, a.w1,
                                            a.gap_open, a.gap_extend);
    }
}

template <bool PROFILE, bool EXACT, bool THRESH>
void launch_variant(const GapblockArgs& a, size_t lds_bytes, hipStream_t s) {
    const uint64_t nch = a.refs.nchunks;
    const uint64_t max_q = (uint64_t(1) << 30) / nch;  // keep a launch under 2^30 workgroups
    for (uint64_t q0 = 0; q0 < a.n_queries; q0 += max_q) {
        const uint64_t nq = a.n_queries - q0 < max_q ? a.n_queries - q0 : max_q;
        hipLaunchKernelGGL((gapblock_kernel<PROFILE, EXACT, THRESH>), dim3(uint32_t(nq * nch)),
                           dim3(kThreads), lds_bytes, s, a, q0);
    }
}

template <bool THRESH>
void launch_mode(const GapblockArgs& a, bool use_profile, size_t lds_bytes, bool exact_wrap,
                 hipStream_t s) {
    if (use_profile) {
        if (exact_wrap) launch_variant<true, true, THRESH>(a, lds_bytes, s);
        else launch_variant<true, false, THRESH>(a, lds_bytes, s);
    } else {
        if (exact_wrap) launch_variant<false, true, THRESH>(a, lds_bytes, s);
        else launch_variant<false, false, THRESH>(a, lds_bytes, s);
    }
}

}  // namespace

void launch_gapblock(const GapblockArgs& a, bool use_profile, uint32_t q_max_len, bool exact_wrap,
                     hipStream_t s) {
    if (a.n_queries == 0 || a.refs.nchunks == 0) return;
    const size_t lds = use_profile ? size_t(q_max_len) * a.A * sizeof(int32_t)
                                   : size_t(a.A) * a.A * sizeof(int32_t);
    const size_t lds_bytes = lds ? lds : sizeof(int32_t);
    if (a.hits.count) launch_mode<true>(a, use_profile, lds_bytes, exact_wrap, s);
    else launch_mode<false>(a, use_profile, lds_bytes, exact_wrap, s);
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
