// Per-pair gap-block scoring shared by the dense / threshold / pair-list kernels
// (gapblock.hip) and the ragged multi-job kernel (batch.hip).
// Reference: src/gapblock.cpp:39-62 (`cell`).
#pragma once

#include "st_internal.hpp"

#include <climits>

namespace st {
namespace gb {

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

// Fast path for a wave whose lanes all hold references of one length n (the common case:
// references are length-sorted inside a chunk and the query is the workgroup's). Then L, d,
// the branch and the prior row are wave-uniform: the positions are fully unrolled under
// scalar guards, prior values come through scalar loads, and with prof (pen) and nprof (-pen)
// in LDS each block position costs one v_add3 (D += p0 - pd), one add (-Sd), one add (prior)
// and one min, plus the two lookups' address arithmetic. Valid only when no int32 sum can
// wrap (the non-exact case) and L <= kFastMaxL. rowoff: A * 4 bytes per profile row.
constexpr uint32_t kFastMaxL = 32;

__device__ __forceinline__ int32_t lds_at(const int32_t* base, uint32_t byte_off) {
    return *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(base) + byte_off);
}

template <class Words>
__device__ __forceinline__ int32_t pair_score_uniform(const int32_t* prof, const int32_t* nprof,
                                                      uint32_t A, const Words& words, uint32_t m,
                                                      uint32_t n, const int32_t* prior,
                                                      uint32_t w1, int32_t go, int32_t ge) {
    const uint32_t L = m < n ? m : n;
    const uint32_t d = m > n ? m - n : n - m;
    const bool q_longer = m >= n;
    const uint32_t row = A * 4;
    const int32_t* __restrict__ prow =
        (prior && d) ? prior + (uint64_t(q_longer ? m : n) * w1 + d) * w1 : nullptr;
    uint32_t D = 0, nS = 0;  // D = P0 - Pd, nS = -Pd
    int32_t best = prow ? prow[0] : 0;
    // byte offset of the second lookup's row relative to the first (diagonal d)
    const uint32_t shift_rows = q_longer ? d * row : 0u;
#pragma unroll
    for (uint32_t w = 0; w < kFastMaxL / 4; ++w) {
        if (4 * w >= L) break;
        const uint32_t x = words(w);
        uint32_t y = x;
        if (!q_longer) {
            const uint32_t t0 = 4 * w + d;
            y = __builtin_amdgcn_alignbyte(words((t0 >> 2) + 1), words(t0 >> 2), t0 & 3u);
        }
#pragma unroll
        for (uint32_t kk = 0; kk < 4; ++kk) {
            const uint32_t t = 4 * w + kk;
            if (t >= L) break;
            const uint32_t a0 = ((x >> (8 * kk)) & 0xFFu) * 4 + t * row;
            if (d == 0) {
                nS -= uint32_t(lds_at(prof, a0));
            } else {
                const uint32_t a1 =
                    q_longer ? a0 + shift_rows : ((y >> (8 * kk)) & 0xFFu) * 4 + t * row;
                const uint32_t p0 = uint32_t(lds_at(prof, a0));
                const uint32_t npd = uint32_t(lds_at(nprof, a1));
                D += p0 + npd;
                nS += npd;
                const int32_t cand = int32_t(D + uint32_t(prow ? prow[t + 1] : 0));
                best = cand < best ? cand : best;
            }
        }
    }
    const uint32_t Sd = 0u - nS;
    return d == 0 ? int32_t(Sd)
                  : int32_t(Sd + uint32_t(best) + uint32_t(go) + (d - 1) * uint32_t(ge));
}

}  // namespace gb
}  // namespace st
