// Per-pair gap-block scoring shared by the dense / threshold / pair-list kernels
// (gapblock.hip) and the ragged multi-job kernel (batch.hip).
// Reference: src/gapblock.cpp:39-62 (`cell`).
//
// Table layout in LDS. A "row" is 64 int32 (256 bytes): entries 0..31 hold pen[a][b] for
// b < A (zero past A), entries 32..63 hold the negations -pen[a][b]. In profile mode row t is
// the query position t (a = q[t]); in table mode row a is the query residue a. Residue codes
// travel as byte offsets (code * 4, at most 124 for A <= 32), so a lookup is one ds_read at
// row * 256 + code4 (+ 128 for the negation), with the row part folded into the
// instruction's immediate offset once positions are unrolled.
#pragma once

#include "st_internal.hpp"

#include <climits>

namespace st {
namespace gb {

constexpr uint32_t kRowInts = 64;
constexpr uint32_t kRowBytes = kRowInts * 4;
constexpr uint32_t kNegBytes = 128;

// Fast path limit: at most this many block-start positions (L = min(m, n)).
constexpr uint32_t kFastMaxL = 32;
// Ints of zero padding after every device copy of a prior cube, so the fast path can read
// kFastMaxL + 1 consecutive entries of any prior row without a bounds check.
constexpr uint32_t kPriorPad = kFastMaxL + 1;

__device__ __forceinline__ int32_t lds_at(const int32_t* base, uint32_t byte_off) {
    return *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Fills `rows` table rows (see above) from the A x A penalty table: row r < valid uses
// residue code_of(r), rows past `valid` are all zero. Called by all threads of a workgroup
// of `threads` threads.
template <class CodeOf>
__device__ __forceinline__ void build_rows(int32_t* lds, uint32_t rows, uint32_t valid,
                                           const int32_t* pen, uint32_t A, CodeOf&& code_of,
                                           uint32_t tid, uint32_t threads) {
    for (uint32_t e = tid; e < rows * kRowInts; e += threads) {
        const uint32_t r = e / kRowInts, c = e % kRowInts, b = c % 32;
        const int32_t v = (r < valid && b < A) ? pen[code_of(r) * A + b] : 0;
        lds[e] = c < 32 ? v : int32_t(0u - uint32_t(v));
    }
}

// ---- short-sequence launches (every query <= kFastMaxL, references <= kShortMaxRef) ------
// The profile gets kFastMaxL zero rows past the query, the alphabet leaves column 31 zero
// (A <= 31) and reference words are padded with code 31 (byte offset kPadByte), so every
// lookup past a pair's own block positions reads 0. The per-query prior slice
// slice[n][i] (row n = reference length, stride kSliceStride) holds the prior of candidate
// i for this query against a length-n reference, or `big` past i = min(m, n); big is chosen
// on the host so that D + big neither wraps nor undercuts a real candidate. Together they
// let a wave whose lanes hold references of different lengths run one straight-line
// sequence of 4 * NW positions.
constexpr uint32_t kShortMaxRef = 64;
constexpr uint32_t kSliceStride = kFastMaxL + 1;
constexpr uint32_t kPadByte = 31u * 4u;

__device__ __forceinline__ void build_slice(int32_t* slice, uint32_t r_max, uint32_t m,
                                            const int32_t* prior, uint32_t w1, int32_t big,
                                            uint32_t tid, uint32_t threads) {
    for (uint32_t e = tid; e < (r_max + 1) * kSliceStride; e += threads) {
        const uint32_t n = e / kSliceStride, i = e % kSliceStride;
        const uint32_t L = m < n ? m : n, d = m > n ? m - n : n - m;
        int32_t v = big;
        if (i <= L) v = (prior && d) ? prior[(uint64_t(m > n ? m : n) * w1 + d) * w1 + i] : 0;
        slice[e] = v;
    }
}

template <bool PROFILE>
struct Look {
    const int32_t* lds;
    const uint8_t* q;
    // positive penalty of query position t against reference byte offset rc4
    __device__ __forceinline__ int32_t operator()(uint32_t t, uint32_t rc4) const {
        if constexpr (PROFILE) {
            return lds_at(lds, t * kRowBytes + rc4);
        } else {
            return lds_at(lds, uint32_t(q[t]) * kRowBytes + rc4);
        }
    }
};

// Global-memory pointer type. Pointers that reach a kernel through a descriptor in memory
// (the ragged batch's jobs) are otherwise generic, and generic (flat) loads count against the
// LDS wait counter as well, serialising them with the penalty lookups.
template <class T>
using GlobalPtr = const __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ GlobalPtr<T> global_ptr(const T* p) {
    return (GlobalPtr<T>)p;
}

// Reference residues as little-endian words of byte offsets: residues 4w..4w+3.
struct PackedWords {  // the chunk-packed layout, one slot
    GlobalPtr<uint32_t> wp;
    uint32_t npad;
    __device__ __forceinline__ uint32_t operator()(uint32_t w) const {
        return wp[uint64_t(w) * npad];
    }
    // the same slot's words from word `w` on
    __device__ __forceinline__ PackedWords shift(uint32_t w) const {
        return {wp + uint64_t(w) * npad, npad};
    }
};

// The chunk-packed layout through a buffer resource over the whole words array (short
// launches; the host keeps the array under 2^31 bytes): the lane's word w is at byte
// voff + w * stride with stride = npad * 4, so the word index goes into the load's scalar
// offset and a load costs no vector address arithmetic.
struct BufferWords {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t voff;
    uint32_t stride;
    __device__ __forceinline__ uint32_t operator()(uint32_t w) const {
        return uint32_t(__builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, w * stride, 0));
    }
    __device__ __forceinline__ BufferWords shift(uint32_t w) const {
        return {rsrc, voff + w * stride, stride};
    }
};

// Buffer resource word 3 for raw (untyped) loads on gfx9-family parts.
constexpr int kBufferFlags = 0x00020000;
constexpr uint64_t kBufferMaxBytes = (uint64_t(1) << 31) - 1;

__device__ __forceinline__ BufferWords buffer_words(const uint32_t* words, uint32_t npad,
                                                    uint32_t nwords, uint32_t slot) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(words), 0, int(uint64_t(nwords + 1) * npad * 4), kBufferFlags);
    return {r, slot * 4u, npad * 4u};
}

struct CsrWords {  // straight from the CSR codes (pair-list mode)
    const uint8_t* r;
    uint32_t n;
    __device__ __forceinline__ uint32_t operator()(uint32_t w) const {
        uint32_t word = 0;
        const uint32_t p = 4 * w;
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b)
            if (p + b < n) word |= (uint32_t(r[p + b]) * 4u) << (8 * b);
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
                const uint32_t rc4 = (x >> (8 * kk)) & 0xFFu;
                const int32_t p0 = look(t, rc4);
                const int32_t pd =
                    q_longer ? look(t + d, rc4) : look(t, (y >> (8 * kk)) & 0xFFu);
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
// references are length-sorted inside a chunk and the query is the workgroup's profile).
// Then L, d and the prior row are wave-uniform, so the pair runs as straight-line code,
// instantiated per word count NW = ceil(L / 4) so every array below has a compile-time size
// and lives in registers: all reference words are loaded up front, the prior row is read
// into scalar registers in one go (device prior cubes carry kPriorPad ints of padding, so
// reading 4 * NW + 1 <= kFastMaxL + 1 entries of any row stays in bounds), and each block
// position costs two ds_reads with immediate row offsets plus a few adds and one min.
// MODE 0: d == 0; MODE 1: query longer; MODE 2: reference longer. Valid only when no int32
// sum can wrap (the non-exact case) and L <= kFastMaxL.
template <int MODE, uint32_t NW, class Words>
__device__ __forceinline__ int32_t pair_score_uniform(const int32_t* prof, const Words& words,
                                                      uint32_t L, uint32_t d,
                                                      GlobalPtr<int32_t> prow, int32_t go,
                                                      int32_t ge) {
    constexpr uint32_t kT = 4 * NW;
    uint32_t x[NW];
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) x[w] = words(w);
    if constexpr (MODE == 0) {
        uint32_t S = 0;
#pragma unroll
        for (uint32_t t = 0; t < kT; ++t) {
            if (t + 4 >= kT && t >= L) break;  // only the last word can be partial
            S += uint32_t(lds_at(prof, t * kRowBytes + ((x[t / 4] >> (8 * (t % 4))) & 0xFFu)));
        }
        return int32_t(S);
    } else {
        uint32_t y[NW];
        if constexpr (MODE == 2) {
            // reference longer: the diagonal-d residue of position t is r[t + d]
            const uint32_t dq = d >> 2, sh = d & 3u;
            const auto wz = words.shift(dq);
            uint32_t z[NW + 1];
#pragma unroll
            for (uint32_t j = 0; j <= NW; ++j) z[j] = wz(j);
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) y[w] = __builtin_amdgcn_alignbyte(z[w + 1], z[w], sh);
        }
        int32_t pv[kT + 1];
        if (prow) {
#pragma unroll
            for (uint32_t t = 0; t <= kT; ++t) pv[t] = prow[t];
        } else {
#pragma unroll
            for (uint32_t t = 0; t <= kT; ++t) pv[t] = 0;
        }
        // MODE 1: the diagonal-d lookup reads query row t + d of the negated half
        const uint32_t off1 = MODE == 1 ? d * kRowBytes + kNegBytes : kNegBytes;
        uint32_t D = 0, nS = 0;  // D = P0 - Pd, nS = -Pd
        int32_t best = pv[0];
#pragma unroll
        for (uint32_t t = 0; t < kT; ++t) {
            if (t + 4 >= kT && t >= L) break;
            const uint32_t rc4 = (x[t / 4] >> (8 * (t % 4))) & 0xFFu;
            const uint32_t p0 = uint32_t(lds_at(prof, t * kRowBytes + rc4));
            uint32_t npd;
            if constexpr (MODE == 1) {
                npd = uint32_t(lds_at(prof, t * kRowBytes + (rc4 + off1)));
            } else {
                const uint32_t rd4 = (y[t / 4] >> (8 * (t % 4))) & 0xFFu;
                npd = uint32_t(lds_at(prof, t * kRowBytes + kNegBytes + rd4));
            }
            D += p0 + npd;
            nS += npd;
            const int32_t cand = int32_t(D + uint32_t(pv[t + 1]));
            best = cand < best ? cand : best;
        }
        const uint32_t Sd = 0u - nS;
        return int32_t(Sd + uint32_t(best) + uint32_t(go) + (d - 1) * uint32_t(ge));
    }
}

template <int MODE, class Words>
__device__ __forceinline__ int32_t pair_score_words(const int32_t* prof, const Words& words,
                                                    uint32_t L, uint32_t d,
                                                    GlobalPtr<int32_t> prow,
                                                    int32_t go, int32_t ge) {
    switch ((L + 3) / 4) {
        case 0:
        case 1: return pair_score_uniform<MODE, 1>(prof, words, L, d, prow, go, ge);
        case 2: return pair_score_uniform<MODE, 2>(prof, words, L, d, prow, go, ge);
        case 3: return pair_score_uniform<MODE, 3>(prof, words, L, d, prow, go, ge);
        case 4: return pair_score_uniform<MODE, 4>(prof, words, L, d, prow, go, ge);
        case 5: return pair_score_uniform<MODE, 5>(prof, words, L, d, prow, go, ge);
        case 6: return pair_score_uniform<MODE, 6>(prof, words, L, d, prow, go, ge);
        case 7: return pair_score_uniform<MODE, 7>(prof, words, L, d, prow, go, ge);
        default: return pair_score_uniform<MODE, 8>(prof, words, L, d, prow, go, ge);
    }
}

// Dispatches the fast path on the (wave-uniform) lengths; requires min(m, n) <= kFastMaxL.
template <class Words>
__device__ __forceinline__ int32_t pair_score_fast(const int32_t* prof, const Words& words,
                                                   uint32_t m, uint32_t n,
                                                   const int32_t* prior, uint32_t w1,
                                                   int32_t go, int32_t ge) {
    const uint32_t L = m < n ? m : n;
    const uint32_t d = m > n ? m - n : n - m;
    if (d == 0) return pair_score_words<0>(prof, words, L, 0, nullptr, go, ge);
    const GlobalPtr<int32_t> prow =
        prior ? global_ptr(prior) + (uint64_t(m > n ? m : n) * w1 + d) * w1 : nullptr;
    if (m > n) return pair_score_words<1>(prof, words, L, d, prow, go, ge);
    return pair_score_words<2>(prof, words, L, d, prow, go, ge);
}

// Mixed-length wave in a short launch: per-lane n (and so L, d, the diagonal and the prior
// row) but one instruction stream of 4 * NW positions, NW = ceil(max_lane L / 4). The
// diagonal-d residue comes from the lane's own realigned words (query longer: the same
// residue, read from row t + d of the negated half; reference longer: r[t + d], gathered
// through per-lane word indices). Positions past the lane's L read zero penalties and
// prior `big`, so they change neither the sums nor the minimum.
template <uint32_t NW, class Words>
__device__ __forceinline__ int32_t pair_score_mixed(const int32_t* prof, const int32_t* slice,
                                                    const Words& words, uint32_t m, uint32_t n,
                                                    int32_t go, int32_t ge) {
    const uint32_t d = m > n ? m - n : n - m;
    const bool ql = m >= n;
    const uint32_t dq = ql ? 0u : d >> 2, sh = ql ? 0u : d & 3u;
    const uint32_t off1 = (ql ? d * kRowBytes : 0u) + kNegBytes;
    const int32_t* prow = slice + n * kSliceStride;
    uint32_t x[NW], z[NW + 1];
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) x[w] = words(w);
    const auto wz = words.shift(dq);
#pragma unroll
    for (uint32_t j = 0; j <= NW; ++j) z[j] = wz(j);
    uint32_t D = 0, nS = 0;  // D = P0 - Pd, nS = -Pd
    int32_t best = prow[0];
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
        const uint32_t y = __builtin_amdgcn_alignbyte(z[w + 1], z[w], sh);
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t t = 4 * w + b;
            const uint32_t rc4 = (x[w] >> (8 * b)) & 0xFFu;
            const uint32_t rd4 = (y >> (8 * b)) & 0xFFu;
            const uint32_t p0 = uint32_t(lds_at(prof, t * kRowBytes + rc4));
            const uint32_t npd = uint32_t(lds_at(prof, t * kRowBytes + (rd4 + off1)));
            D += p0 + npd;
            nS += npd;
            const int32_t cand = int32_t(D + uint32_t(prow[t + 1]));
            best = cand < best ? cand : best;
        }
        // keep at most one word of lookups in flight: bounds the registers (and so keeps
        // occupancy up) at the cost of per-word LDS latency, which other waves hide
        __builtin_amdgcn_sched_barrier(0);
    }
    const uint32_t Sd = 0u - nS;
    return d == 0 ? int32_t(Sd)
                  : int32_t(Sd + uint32_t(best) + uint32_t(go) + (d - 1) * uint32_t(ge));
}

// One reference slot of a short launch: the uniform fast path when the wave's references
// share one length, else the mixed path. m is the workgroup's query length (<= kFastMaxL).
// The wave must hold 64 consecutive slots of the chunk-packed layout.
template <class Words>
__device__ __forceinline__ int32_t pair_score_short(const int32_t* prof, const int32_t* slice,
                                                    const Words& words, uint32_t k, uint32_t m,
                                                    uint32_t n, const int32_t* prior,
                                                    uint32_t w1, int32_t go, int32_t ge) {
    const uint32_t n0 = __builtin_amdgcn_readfirstlane(n);
    if (__all(k == kNoSlot || n == n0)) return pair_score_fast(prof, words, m, n0, prior, w1, go, ge);
    // The wave's longest reference: a wave holds 64 consecutive slots of a length-sorted
    // chunk whose padding slots (length 0) come last, so it is the length in its last real
    // slot (and some lane is real here: an all-padding wave takes the uniform path).
    const uint64_t real = __ballot(k != kNoSlot);
    const uint32_t last = 63u - uint32_t(__builtin_clzll(real));
    const uint32_t nmax = uint32_t(__builtin_amdgcn_readlane(int(n), int(last)));
    const uint32_t Lmax = m < nmax ? m : nmax;
    switch ((Lmax + 3) / 4) {
        case 0:
        case 1: return pair_score_mixed<1>(prof, slice, words, m, n, go, ge);
        case 2: return pair_score_mixed<2>(prof, slice, words, m, n, go, ge);
        case 3: return pair_score_mixed<3>(prof, slice, words, m, n, go, ge);
        case 4: return pair_score_mixed<4>(prof, slice, words, m, n, go, ge);
        case 5: return pair_score_mixed<5>(prof, slice, words, m, n, go, ge);
        case 6: return pair_score_mixed<6>(prof, slice, words, m, n, go, ge);
        case 7: return pair_score_mixed<7>(prof, slice, words, m, n, go, ge);
        default: return pair_score_mixed<8>(prof, slice, words, m, n, go, ge);
    }
}

// ---- twin rows: two queries of one length in the 16-bit halves of every word ------------
// Twin tables hold no negations: a twin row is 32 words (kTwinRowBytes), and the diagonal
// penalty is subtracted with a packed 16-bit subtract (the same bits as adding the stored
// 16-bit negation), which halves the profile's LDS and so raises the workgroups per CU.
// Short launches can score two queries a and b of the same length m at once: every table
// word holds query a's value in its low half and query b's in its high half (profile rows
// pen[q_a[t]][c] | pen[q_b[t]][c] << 16 and the negations; prior-slice entries duplicated).
// With equal lengths the two pairs share d, the diagonal and the prior row, so one lookup,
// one address and one packed add or min serve both. The host admits a twin launch only when
// 6 T + R < 2^15 (T = q_max * max|pen|, R = max|prior|, the short-launch bound at 16 bits):
// then |D| <= 2T, |nS| <= T, real candidates lie in [-2T - R, 2T + R], the sentinel
// big = 2^15 - 1 - 2T keeps D + big in (2T + R, 2^15 - 1], and no half ever wraps, so the
// signed 16-bit minimum and the final int32 combination equal the int32 path bit for bit.
constexpr uint32_t kTwinRowInts = 32;
constexpr uint32_t kTwinRowBytes = kTwinRowInts * 4;

using u16x2 = uint16_t __attribute__((ext_vector_type(2)));
using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
using i16x2 = int16_t __attribute__((ext_vector_type(2)));

// An empty asm statement that takes and returns v in a VGPR: the compiler must have v's value
// at this point, so it cannot re-associate the arithmetic that produces v across it.
template <class T>
__device__ __forceinline__ void pin(T& v) {
    static_assert(sizeof(T) == 4, "pin takes one 32-bit value");
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    asm volatile("" : "+v"(u));
    v = __builtin_bit_cast(T, u);
}

__device__ __forceinline__ uint32_t pack16(int32_t lo, int32_t hi) {
    return (uint32_t(lo) & 0xFFFFu) | (uint32_t(hi) << 16);
}

__device__ __forceinline__ u16x2 lds_at2(const uint32_t* base, uint32_t byte_off) {
    return __builtin_bit_cast(
        u16x2, *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + byte_off));
}

// Twin profile rows for two queries of length `valid` (codes qa, qb): row r entry b is
// pen[qa[r]][b] | pen[qb[r]][b] << 16 for r < valid and b < A, zero elsewhere.
__device__ __forceinline__ void build_rows2(uint32_t* lds, uint32_t rows, uint32_t valid,
                                            const int32_t* pen, uint32_t A, const uint8_t* qa,
                                            const uint8_t* qb, uint32_t tid, uint32_t threads) {
    for (uint32_t e = tid; e < rows * kTwinRowInts; e += threads) {
        const uint32_t r = e / kTwinRowInts, b = e % kTwinRowInts;
        int32_t va = 0, vb = 0;
        if (r < valid && b < A) {
            va = pen[uint32_t(qa[r]) * A + b];
            vb = pen[uint32_t(qb[r]) * A + b];
        }
        lds[e] = pack16(va, vb);
    }
}

// The twin prior slice: build_slice's entries as single 16-bit values (both queries of a twin
// share them), kTwinSliceHalves to a row so that every row starts 16-byte aligned and a lane
// reads eight entries with one 16-byte load. Entries past the candidates (i > min(m, n),
// up to the row's end) hold big, the 16-bit sentinel.
constexpr uint32_t kTwinSliceHalves = 40;
constexpr uint32_t kTwinSliceBytes = kTwinSliceHalves * 2;

__device__ __forceinline__ void build_slice2(uint32_t* slice, uint32_t r_max, uint32_t m,
                                             const int32_t* prior, uint32_t w1, int32_t big,
                                             uint32_t tid, uint32_t threads) {
    uint16_t* s16 = reinterpret_cast<uint16_t*>(slice);
    for (uint32_t e = tid; e < (r_max + 1) * kTwinSliceHalves; e += threads) {
        const uint32_t n = e / kTwinSliceHalves, i = e % kTwinSliceHalves;
        const uint32_t L = m < n ? m : n, d = m > n ? m - n : n - m;
        int32_t v = big;
        if (i <= L) v = (prior && d) ? prior[(uint64_t(m > n ? m : n) * w1 + d) * w1 + i] : 0;
        s16[e] = uint16_t(uint32_t(v));
    }
}

// Entry i of a twin slice row held as 16-byte groups g[i / 8], in both halves of a word (i is
// a constant once the caller's loops are unrolled).
__device__ __forceinline__ u16x2 twin_prior(const u32x4* g, uint32_t i) {
    // the element goes through a scalar first: __builtin_bit_cast applied to a vector
    // element expression directly reads element 0 with this compiler
    const uint32_t e = g[i / 8][(i % 8) / 2];
    const u16x2 w = __builtin_bit_cast(u16x2, e);
    return (i & 1u) ? __builtin_shufflevector(w, w, 1, 1) : __builtin_shufflevector(w, w, 0, 0);
}

// The int32 result of one half: Sd = -nS, then Sd + best + go + (d - 1) ge (d > 0), in the
// wrapping order of pair_score_mixed.
__device__ __forceinline__ int32_t twin_finish(int16_t ns, int16_t best, uint32_t d, int32_t go,
                                               int32_t ge) {
    const uint32_t Sd = 0u - uint32_t(int32_t(ns));
    return d == 0 ? int32_t(Sd)
                  : int32_t(Sd + uint32_t(int32_t(best)) + uint32_t(go) + (d - 1) * uint32_t(ge));
}

// pair_score_mixed on twin rows: per-lane n, one instruction stream of 4 * NW positions;
// x: the reference words from position 0 (NW of them), z: the words from the lane's
// diagonal word offset (NW + 1), both already in registers.
template <uint32_t NW>
__device__ __forceinline__ void twin_mixed_regs(const uint32_t* prof, const uint32_t* slice,
                                                const uint32_t* x, const uint32_t* z,
                                                uint32_t m, uint32_t n, int32_t go, int32_t ge,
                                                int32_t& sa, int32_t& sb) {
    const uint32_t d = m > n ? m - n : n - m;
    const bool ql = m >= n;
    const uint32_t sh = ql ? 0u : d & 3u;
    const uint32_t off1 = ql ? d * kTwinRowBytes : 0u;
    // the lane's prior row as 16-byte groups of eight entries; word w uses entries 4w + 1 ..
    // 4w + 4, so group g is read at the first word that needs it (w = 0, then every odd w)
    const u32x4* prow = reinterpret_cast<const u32x4*>(
        reinterpret_cast<const char*>(slice) + n * kTwinSliceBytes);
    constexpr uint32_t NG = (4 * NW + 1 + 7) / 8;
    u32x4 g[NG];
    g[0] = prow[0];
    u16x2 D = {0, 0}, nS = {0, 0};  // per half: D = P0 - Pd, nS = -Pd
    i16x2 best = __builtin_bit_cast(i16x2, twin_prior(g, 0));
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
        if (w % 2 == 1 && (w + 1) / 2 < NG) g[(w + 1) / 2] = prow[(w + 1) / 2];
        const uint32_t y = __builtin_amdgcn_alignbyte(z[w + 1], z[w], sh);
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t t = 4 * w + b;
            const uint32_t rc4 = (x[w] >> (8 * b)) & 0xFFu;
            const uint32_t rd4 = (y >> (8 * b)) & 0xFFu;
            const u16x2 p0 = lds_at2(prof, t * kTwinRowBytes + rc4);
            const u16x2 pd = lds_at2(prof, t * kTwinRowBytes + (rd4 + off1));
            D += p0;
            D -= pd;
            nS -= pd;
            const i16x2 cand = __builtin_bit_cast(i16x2, D + twin_prior(g, t + 1));
            best = __builtin_elementwise_min(best, cand);
        }
        // the running sums and minimum are materialised once per word: no reassociation
        // keeps a word's lookups live past it
        pin(D);
        pin(nS);
        pin(best);
        __builtin_amdgcn_sched_barrier(0);
    }
    const i16x2 ns = __builtin_bit_cast(i16x2, nS);
    sa = twin_finish(ns.x, best.x, d, go, ge);
    sb = twin_finish(ns.y, best.y, d, go, ge);
}

// A lane's words for the scorers: x = words 0..NX-1, z = NZ words from its diagonal word
// offset (d / 4 when the reference is longer, else 0). When no lane's reference is longer
// than the query every offset is 0 and z repeats x, so only the words past x are loaded.
// Every lane of the wave must be active.
template <uint32_t NX, uint32_t NZ, class Words>
__device__ __forceinline__ void load_words(const Words& words, uint32_t m, uint32_t n,
                                           uint32_t* x, uint32_t* z) {
#pragma unroll
    for (uint32_t w = 0; w < NX; ++w) x[w] = words(w);
    if (__all(m >= n)) {
#pragma unroll
        for (uint32_t j = 0; j < NZ; ++j) z[j] = j < NX ? x[j] : words(j);
    } else {
        const auto wz = words.shift(m >= n ? 0u : (n - m) >> 2);
#pragma unroll
        for (uint32_t j = 0; j < NZ; ++j) z[j] = wz(j);
    }
}

// pair_score_mixed on twin rows: per-lane n, one instruction stream of 4 * NW positions.
template <uint32_t NW, class Words>
__device__ __forceinline__ void twin_score_mixed(const uint32_t* prof, const uint32_t* slice,
                                                 const Words& words, uint32_t m, uint32_t n,
                                                 int32_t go, int32_t ge, int32_t& sa,
                                                 int32_t& sb) {
    uint32_t x[NW], z[NW + 1];
    load_words<NW, NW + 1>(words, m, n, x, z);
    twin_mixed_regs<NW>(prof, slice, x, z, m, n, go, ge, sa, sb);
}

// Every real lane has n == m (d == 0): the two sums only, over the m (wave-uniform) positions.
template <uint32_t NW>
__device__ __forceinline__ void twin_sum_regs(const uint32_t* prof, const uint32_t* x,
                                              uint32_t m, int32_t& sa, int32_t& sb) {
    u16x2 S = {0, 0};
#pragma unroll
    for (uint32_t t = 0; t < 4 * NW; ++t) {
        if (t + 4 >= 4 * NW && t >= m) break;  // only the last word can be partial
        S += lds_at2(prof, t * kTwinRowBytes + ((x[t / 4] >> (8 * (t % 4))) & 0xFFu));
    }
    const i16x2 s = __builtin_bit_cast(i16x2, S);
    sa = s.x;
    sb = s.y;
}

template <uint32_t NW, class Words>
__device__ __forceinline__ void twin_score_sum(const uint32_t* prof, const Words& words,
                                               uint32_t m, int32_t& sa, int32_t& sb) {
    uint32_t x[NW];
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) x[w] = words(w);
    twin_sum_regs<NW>(prof, x, m, sa, sb);
}

// pair_score_short on twin rows: both scores of reference slot k against the twin queries
// (length m <= kFastMaxL). The wave must hold 64 consecutive slots of the chunk-packed layout.
template <class Words>
__device__ __forceinline__ void twin_score_short(const uint32_t* prof, const uint32_t* slice,
                                                 const Words& words, uint32_t k, uint32_t m,
                                                 uint32_t n, int32_t go, int32_t ge, int32_t& sa,
                                                 int32_t& sb) {
    const uint64_t real = __ballot(k != kNoSlot);
    if (real == 0) {  // an all-padding wave
        sa = sb = 0;
        return;
    }
    if (__all(k == kNoSlot || n == m)) {
        switch ((m + 3) / 4) {
            case 0:
            case 1: return twin_score_sum<1>(prof, words, m, sa, sb);
            case 2: return twin_score_sum<2>(prof, words, m, sa, sb);
            case 3: return twin_score_sum<3>(prof, words, m, sa, sb);
            case 4: return twin_score_sum<4>(prof, words, m, sa, sb);
            case 5: return twin_score_sum<5>(prof, words, m, sa, sb);
            case 6: return twin_score_sum<6>(prof, words, m, sa, sb);
            case 7: return twin_score_sum<7>(prof, words, m, sa, sb);
            default: return twin_score_sum<8>(prof, words, m, sa, sb);
        }
    }
    // the wave's longest reference is the length in its last real slot (as pair_score_short)
    const uint32_t last = 63u - uint32_t(__builtin_clzll(real));
    const uint32_t nmax = uint32_t(__builtin_amdgcn_readlane(int(n), int(last)));
    const uint32_t Lmax = m < nmax ? m : nmax;
    switch ((Lmax + 3) / 4) {
        case 0:
        case 1: return twin_score_mixed<1>(prof, slice, words, m, n, go, ge, sa, sb);
        case 2: return twin_score_mixed<2>(prof, slice, words, m, n, go, ge, sa, sb);
        case 3: return twin_score_mixed<3>(prof, slice, words, m, n, go, ge, sa, sb);
        case 4: return twin_score_mixed<4>(prof, slice, words, m, n, go, ge, sa, sb);
        case 5: return twin_score_mixed<5>(prof, slice, words, m, n, go, ge, sa, sb);
        case 6: return twin_score_mixed<6>(prof, slice, words, m, n, go, ge, sa, sb);
        case 7: return twin_score_mixed<7>(prof, slice, words, m, n, go, ge, sa, sb);
        default: return twin_score_mixed<8>(prof, slice, words, m, n, go, ge, sa, sb);
    }
}

}  // namespace gb
}  // namespace st
