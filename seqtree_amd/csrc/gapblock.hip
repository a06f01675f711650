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

// Visits block-start positions t = 0..L-1 in order, handing the two penalties to f.
template <bool PROFILE, class F>
__device__ __forceinline__ void walk(const Look<PROFILE>& look, const uint32_t* __restrict__ wp,
                                     uint32_t npad, uint32_t L, uint32_t d, bool q_longer, F&& f) {
    for (uint32_t w = 0; 4 * w < L; ++w) {
        const uint32_t x = wp[uint64_t(w) * npad];
        uint32_t y = 0;
        if (!q_longer) {
            const uint32_t t0 = 4 * w + d;
            const uint32_t lo = wp[uint64_t(t0 >> 2) * npad];
            const uint32_t hi = wp[uint64_t((t0 >> 2) + 1) * npad];
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

template <bool PROFILE, bool EXACT>
__global__ __launch_bounds__(kThreads) void gapblock_kernel(GapblockArgs a, uint64_t q_base) {
    extern __shared__ int32_t lds[];
    __shared__ int32_t stage[kChunk];

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
    const uint32_t npad = a.refs.npad;
    const uint32_t w1 = a.w1;

    for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
        const uint32_t slot = c * kChunk + it * kThreads + tid;
        const uint32_t k = a.refs.orig[slot];
        if (k == kNoSlot) continue;
        const uint32_t n = a.refs.len[slot];
        const uint32_t L = m < n ? m : n;
        const uint32_t d = m > n ? m - n : n - m;
        const bool q_longer = m >= n;
        const int32_t* prow =
            (a.prior && d) ? a.prior + (uint64_t(q_longer ? m : n) * w1 + d) * w1 : nullptr;
        const uint32_t* wp = a.refs.words + slot;

        int32_t score;
        if constexpr (!EXACT) {
            uint32_t P0 = 0, Pd = 0;
            int32_t best = prow ? prow[0] : 0;
            walk<PROFILE>(look, wp, npad, L, d, q_longer, [&](uint32_t t, int32_t p0, int32_t pd) {
                P0 += uint32_t(p0);
                Pd += uint32_t(pd);
                if (d) {
                    const int32_t cand = int32_t(P0 - Pd + uint32_t(prow ? prow[t + 1] : 0));
                    best = cand < best ? cand : best;
                }
            });
            score = d == 0 ? int32_t(Pd)
                           : int32_t(Pd + uint32_t(best) + uint32_t(a.gap_open) +
                                     (d - 1) * uint32_t(a.gap_extend));
        } else {
            uint32_t Sd = 0;
            walk<PROFILE>(look, wp, npad, L, d, q_longer,
                          [&](uint32_t, int32_t, int32_t pd) { Sd += uint32_t(pd); });
            if (d == 0) {
                score = int32_t(Sd);
            } else {
                uint32_t P0 = 0, Pd = 0;
                int32_t best = INT_MAX;
                // candidate i = t is formed before position t is consumed, i = L after the loop
                walk<PROFILE>(look, wp, npad, L, d, q_longer,
                              [&](uint32_t t, int32_t p0, int32_t pd) {
                                  const uint32_t pr = prow ? uint32_t(prow[t]) : 0u;
                                  const int32_t cand = int32_t((P0 + (Sd - Pd)) + pr);
                                  best = cand < best ? cand : best;
                                  P0 += uint32_t(p0);
                                  Pd += uint32_t(pd);
                              });
                {
                    const uint32_t pr = prow ? uint32_t(prow[L]) : 0u;
                    const int32_t cand = int32_t((P0 + (Sd - Pd)) + pr);
                    best = cand < best ? cand : best;
                }
                score = int32_t(uint32_t(best) + uint32_t(a.gap_open) +
                                (d - 1) * uint32_t(a.gap_extend));
            }
        }
        stage[k - c * kChunk] = score;
    }
    __syncthreads();

    const uint64_t col0 = uint64_t(c) * kChunk;
    const uint64_t lim64 = a.n_refs - col0;
    const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
    int32_t* row = a.out + i * a.n_refs + col0;
    for (uint32_t t = tid; t < lim; t += kThreads) row[t] = stage[t];
}

template <bool PROFILE, bool EXACT>
void launch_variant(const GapblockArgs& a, size_t lds_bytes, hipStream_t s) {
    const uint64_t nch = a.refs.nchunks;
    const uint64_t max_q = (uint64_t(1) << 30) / nch;  // keep a launch under 2^30 workgroups
    for (uint64_t q0 = 0; q0 < a.n_queries; q0 += max_q) {
        const uint64_t nq = a.n_queries - q0 < max_q ? a.n_queries - q0 : max_q;
        hipLaunchKernelGGL((gapblock_kernel<PROFILE, EXACT>), dim3(uint32_t(nq * nch)),
                           dim3(kThreads), lds_bytes, s, a, q0);
    }
}

}  // namespace

void launch_gapblock(const GapblockArgs& a, bool use_profile, uint32_t q_max_len, bool exact_wrap,
                     hipStream_t s) {
    if (a.n_queries == 0 || a.refs.nchunks == 0) return;
    const size_t lds = use_profile ? size_t(q_max_len) * a.A * sizeof(int32_t)
                                   : size_t(a.A) * a.A * sizeof(int32_t);
    const size_t lds_bytes = lds ? lds : sizeof(int32_t);
    if (use_profile) {
        if (exact_wrap) launch_variant<true, true>(a, lds_bytes, s);
        else launch_variant<true, false>(a, lds_bytes, s);
    } else {
        if (exact_wrap) launch_variant<false, true>(a, lds_bytes, s);
        else launch_variant<false, false>(a, lds_bytes, s);
    }
}

}  // namespace st
