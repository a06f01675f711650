// Gotoh affine Needleman-Wunsch / Smith-Waterman scoring kernels (reference:
// src/pairwise.cpp:76-125 `score_encoded`, driven by src/pairwise.cpp:273-334).
//
// Recurrence (maximising; gap costs are positive magnitudes, X<->Y switches pay gap_open):
//   M[i][j] = s(q_i, r_j) + max(M, X, Y)[i-1][j-1]        (local: floored at 0)
//   X[i][j] = max(M[i-1][j] - go, X[i-1][j] - ge, Y[i-1][j] - go)
//   Y[i][j] = max(M[i][j-1] - go, Y[i][j-1] - ge, X[i][j-1] - go)
// Global: X[i][0] = -gap_run(i), Y[0][j] = -gap_run(j), result max(M,X,Y)[m][n].
// Local: M boundary 0, gap states unreachable on the boundary, result max over all M (>= 0).
//
// Register kernel: workgroup = (query, reference chunk); each lane owns one reference and
// sweeps its columns j, keeping a stripe of R query rows in registers (R >= m, R a multiple
// of 4; rows past m are dummies whose profile entries can never win). With U = max(M, Y) and
// V = max(M, X) the cell update is 10 VALU ops (11 local):
//   M = Hdiag + s;  X = max(U_up - go, X_up - ge);  Y = max(V_left - go, Y_left - ge)
//   U = max(M, Y);  V = max(M, X);  H = max(U, X)
// The substitution score comes from a query profile in LDS, prof[row][code] with a fixed
// 32-entry row stride, so the lookup is one ds_read with an immediate row offset.
// When a host-side range proof shows every reachable value, the -2^14 sentinel and one gap
// step below it fit int16, two queries share a workgroup and ride in the two halves of each
// register: the same 10 instructions, as v_pk_* ops, advance two cells, and one profile read
// (both queries' scores packed in one dword) feeds both. With the proof only for int32 the
// kernel runs one query in int32 (-2^30 sentinel); otherwise the int64 generic kernel
// (a straight two-row DP per thread) computes exactly what the reference's int64 DP does,
// including the narrow() overflow check.

#include "st_internal.hpp"

namespace st {
namespace {

constexpr int32_t kNeg32 = -(1 << 30);
constexpr int16_t kNeg16 = -(1 << 14);

typedef short v2i16 __attribute__((ext_vector_type(2)));

// Cell arithmetic of the register kernel. I32: one query per workgroup, int32 cells.
// I16x2: two queries per workgroup packed into the 16-bit halves of every register, so one
// v_pk_* instruction advances two DP cells (same reference column, two query rows).
struct I32 {
    using T = int32_t;
    static constexpr int Q = 1;
    static constexpr int32_t kNeg = kNeg32;
    __device__ static T splat(int32_t v) { return v; }
    __device__ static T vmax(T a, T b) { return max(a, b); }
    __device__ static T load(uint32_t w) { return int32_t(w); }
    __device__ static uint32_t pack(const int32_t* v) { return uint32_t(v[0]); }
    __device__ static int32_t half(T v, int) { return v; }
};

struct I16x2 {
    using T = v2i16;
    static constexpr int Q = 2;
    static constexpr int32_t kNeg = kNeg16;
    __device__ static T splat(int32_t v) { return T{int16_t(v), int16_t(v)}; }
    __device__ static T vmax(T a, T b) { return __builtin_elementwise_max(a, b); }
    __device__ static T load(uint32_t w) { return __builtin_bit_cast(T, w); }
    __device__ static uint32_t pack(const int32_t* v) {
        return (uint32_t(uint16_t(int16_t(v[0])))) | (uint32_t(uint16_t(int16_t(v[1]))) << 16);
    }
    __device__ static int32_t half(T v, int h) { return h == 0 ? int32_t(v.x) : int32_t(v.y); }
};

// Register-stripe Gotoh: workgroup = (C::Q queries of one length class, reference chunk).
//
// HFORM (used when gap_open >= gap_extend): because X - go <= X - ge and Y - go <= Y - ge,
//   X[i][j] = max(H[i-1][j] - go, X[i-1][j] - ge),  Y[i][j] = max(H[i][j-1] - go, Y[i][j-1] - ge)
// with H = max(M, X, Y). Keeping Hg = H - go per cell and folding go into the profile
// (prof = s + go, so M = Hg_diag + prof) leaves 8 instructions per cell (10 local):
//   M = Hgd + s';  X = max(Hg_up, X_up - ge);  Y = max(Hg_left, Y_left - ge);
//   Hg = max(M, X, Y) - go
// and two registers per stripe row (Hg, Y). Otherwise the U/V form above is used.
template <class C, int R, bool LOCAL, bool HFORM>
__global__ __launch_bounds__(kThreads) void gotoh_reg_kernel(GotohArgs a,
                                                             const uint32_t* __restrict__ list,
                                                             uint32_t count, uint64_t base) {
    using T = typename C::T;
    constexpr int Q = C::Q;
    __shared__ uint32_t prof[R * 32];
    __shared__ int32_t stage[Q][kChunk];

    const uint32_t tid = threadIdx.x;
    const uint32_t nch = a.refs.nchunks;
    const uint32_t c = blockIdx.x % nch;
    const uint64_t g = base + blockIdx.x / nch;

    uint32_t qi[Q], m[Q];
    bool valid[Q];
    const uint8_t* q[Q];
#pragma unroll
    for (int h = 0; h < Q; ++h) {
        const uint64_t idx = g * Q + h;
        valid[h] = idx < count;
        qi[h] = list[valid[h] ? idx : g * Q];
        const uint64_t q0 = a.q_off[qi[h]];
        m[h] = uint32_t(a.q_off[qi[h] + 1] - q0);
        q[h] = a.q_codes + q0;
    }
    const uint32_t A = a.A;
    const int32_t shift = HFORM ? a.gap_open : 0;  // profile holds s + go in the H form
    const int32_t dummy = LOCAL ? C::kNeg : shift;  // rows past a query's end never win
    for (uint32_t e = tid; e < uint32_t(R) * 32; e += kThreads) {
        const uint32_t r = e >> 5, b = e & 31u;
        int32_t v[Q];
#pragma unroll
        for (int h = 0; h < Q; ++h)
            v[h] = (r < m[h] && b < A) ? a.sim[uint32_t(q[h][r]) * A + b] + shift : dummy;
        prof[e] = C::pack(v);
    }
    __syncthreads();

    const T go = C::splat(a.gap_open), ge = C::splat(a.gap_extend);
    const T neg = C::splat(C::kNeg), zero = C::splat(0);
    const uint32_t npad = a.refs.npad;

    for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
        const uint32_t slot = c * kChunk + it * kThreads + tid;
        const uint32_t k = a.refs.orig[slot];
        if (k == kNoSlot) continue;
        const uint32_t n = a.refs.len[slot];

        // HFORM: Hs = H - go, Y.  U/V form: V = max(M, X), Y, Hs = H.
        T V[HFORM ? 1 : R], Y[R], Hs[R];
        {
            // column 0: H[i][0] = -gap_run(i) (global) or 0 (local), Y = -inf, X = H
            T gr = LOCAL ? zero : C::splat(-a.gap_open);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                Hs[r] = HFORM ? gr - go : gr;
                if constexpr (!HFORM) V[r] = gr;
                Y[r] = neg;
                if (!LOCAL) gr = gr - ge;
            }
        }
        T hrow0 = HFORM ? zero - go : zero;  // H[0][j] (minus go in the H form), j = 0
        T best = zero;
        const uint32_t* wp = a.refs.words + slot;
        uint32_t word = 0;
        T g0 = C::splat(-a.gap_open);  // -gap_run(j + 1) on row 0
        for (uint32_t j = 0; j < n; ++j) {
            if ((j & 3u) == 0) word = wp[uint64_t(j >> 2) * npad];
            const uint32_t rc = (word >> (8 * (j & 3u))) & 0xFFu;
            const uint32_t* pcol = prof + rc;
            T up, hd;  // H form: Hg_up ; U/V form: U_up
            T xup = neg;
            if (LOCAL) {
                up = HFORM ? zero - go : zero;
                hd = HFORM ? zero - go : zero;
            } else {
                up = HFORM ? g0 - go : g0;
                hd = hrow0;
                hrow0 = up;
                g0 = g0 - ge;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const T s = C::load(pcol[r * 32]);
                T mv = hd + s;
                if (LOCAL) mv = C::vmax(mv, zero);
                if constexpr (HFORM) {
                    const T xv = C::vmax(up, xup - ge);
                    const T yv = C::vmax(Hs[r], Y[r] - ge);
                    hd = Hs[r];
                    Hs[r] = C::vmax(C::vmax(mv, xv), yv) - go;
                    up = Hs[r];
                    xup = xv;
                    Y[r] = yv;
                } else {
                    const T xv = C::vmax(up - go, xup - ge);
                    const T yv = C::vmax(V[r] - go, Y[r] - ge);
                    hd = Hs[r];
                    up = C::vmax(mv, yv);
                    xup = xv;
                    V[r] = C::vmax(mv, xv);
                    Y[r] = yv;
                    Hs[r] = C::vmax(up, xv);
                }
                if (LOCAL) best = C::vmax(best, mv);
            }
        }

#pragma unroll
        for (int h = 0; h < Q; ++h) {
            int32_t res;
            if (LOCAL) {
                res = C::half(best, h);
            } else if (m[h] == 0) {
                res = n == 0 ? 0 : -(a.gap_open + int32_t(n - 1) * a.gap_extend);
            } else {
                res = C::half(Hs[0], h);
#pragma unroll
                for (int r = 1; r < R; ++r)
                    if (uint32_t(r) == m[h] - 1) res = C::half(Hs[r], h);
                if (HFORM) res += a.gap_open;
            }
            if (a.dist) res = int32_t(a.q_self[qi[h]] + a.r_self[k] - 2 * int64_t(res));
            stage[h][k - c * kChunk] = res;
        }
    }
    __syncthreads();

    const uint64_t col0 = uint64_t(c) * kChunk;
    const uint64_t lim64 = a.n_refs - col0;
    const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
#pragma unroll
    for (int h = 0; h < Q; ++h) {
        if (!valid[h]) continue;
        int32_t* row = a.out + uint64_t(qi[h]) * a.n_refs + col0;
        for (uint32_t t = tid; t < lim; t += kThreads) row[t] = stage[h][t];
    }
}

// Multi-stripe register Gotoh for queries longer than one register stripe (germline V genes
// of ~90 residues and beyond): the same lane-per-reference sweep and the same cell update as
// gotoh_reg_kernel, repeated over the query in stripes of R rows. Between stripes each lane
// keeps the bottom row of its stripe -- per reference column the H (H - go in the H form) and
// X values, and U in the U/V form -- in its own lane-interleaved slice of `bnd`, which the
// next stripe's top row reads as "the row above" and overwrites in place, column by column.
// The grid is persistent (units = query group x reference chunk, strided by gridDim.x) so
// the boundary scratch is sized by the grid, not by the matrix; every workgroup owns
// NB * W * kThreads words of it. A query's score is taken from the stripe holding its last
// row; two packed queries of different lengths run to the longer one's last stripe.
//
// prof_rows: the LDS profile's rows (dynamic LDS). When they hold every stripe of the longest
// query (full profile), the profile is built once per unit and the slot and stripe loops run
// without barriers; otherwise (prof_rows == R) each stripe rebuilds its R rows between two
// barriers.
template <class C, int R, bool LOCAL, bool HFORM>
__global__ __launch_bounds__(kThreads) void gotoh_stripes_kernel(
    GotohArgs a, const uint32_t* __restrict__ list, uint32_t count, uint64_t n_units,
    uint32_t* __restrict__ bnd_all, uint32_t W, uint32_t prof_rows) {
    using T = typename C::T;
    constexpr int Q = C::Q;
    constexpr uint32_t NB = HFORM ? 2 : 3;  // boundary words per column: H, X (, U)
    extern __shared__ uint32_t prof[];  // [prof_rows][32]
    __shared__ int32_t stage[Q][kChunk];

    const uint32_t tid = threadIdx.x;
    const uint32_t nch = a.refs.nchunks;
    uint32_t* bnd = bnd_all + uint64_t(blockIdx.x) * NB * W * kThreads + tid;
    const uint32_t A = a.A;
    const int32_t shift = HFORM ? a.gap_open : 0;  // profile holds s + go in the H form
    const int32_t dummy = LOCAL ? C::kNeg : shift;  // rows past a query's end never win
    const T go = C::splat(a.gap_open), ge = C::splat(a.gap_extend);
    const T neg = C::splat(C::kNeg), zero = C::splat(0);
    const uint32_t npad = a.refs.npad;

    for (uint64_t u = blockIdx.x; u < n_units; u += gridDim.x) {
        const uint32_t c = uint32_t(u % nch);
        const uint64_t g = u / nch;
        uint32_t qi[Q], m[Q];
        bool valid[Q];
        const uint8_t* q[Q];
        uint32_t rows = 0;
#pragma unroll
        for (int h = 0; h < Q; ++h) {
            const uint64_t idx = g * Q + h;
            valid[h] = idx < count;
            qi[h] = list[valid[h] ? idx : g * Q];
            const uint64_t q0 = a.q_off[qi[h]];
            m[h] = uint32_t(a.q_off[qi[h] + 1] - q0);
            q[h] = a.q_codes + q0;
            rows = m[h] > rows ? m[h] : rows;
        }
        const uint32_t S = rows == 0 ? 1 : (rows + R - 1) / R;
        const bool full = S * uint32_t(R) <= prof_rows;  // uniform
        auto build = [&](uint32_t r0, uint32_t nr) {  // profile rows r0 .. r0 + nr - 1
            for (uint32_t e = tid; e < nr * 32; e += kThreads) {
                const uint32_t r = e >> 5, b = e & 31u, i = r0 + r;
                int32_t v[Q];
#pragma unroll
                for (int h = 0; h < Q; ++h)
                    v[h] = (i < m[h] && b < A) ? a.sim[uint32_t(q[h][i]) * A + b] + shift
                                               : dummy;
                prof[(full ? i : r) * 32 + b] = C::pack(v);
            }
        };
        if (full) {
            __syncthreads();  // every lane is past the previous unit's profile and stage
            build(0, S * uint32_t(R));
            __syncthreads();
        }

        for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
            const uint32_t slot = c * kChunk + it * kThreads + tid;
            const uint32_t k = a.refs.orig[slot];
            const bool live = k != kNoSlot;
            const uint32_t n = live ? a.refs.len[slot] : 0u;  // padding lanes sweep nothing
            const uint32_t* wp = a.refs.words + slot;
            int32_t res[Q];
#pragma unroll
            for (int h = 0; h < Q; ++h) res[h] = 0;
            T best = zero;

            for (uint32_t sidx = 0; sidx < S; ++sidx) {
                const uint32_t i0 = sidx * R;  // this stripe holds DP rows i0 + 1 .. i0 + R
                if (!full) {
                    __syncthreads();  // every lane is past the previous stripe's profile
                    build(i0, R);
                    __syncthreads();
                }
                const uint32_t* sprof = prof + (full ? i0 * 32 : 0u);

                // column 0: H[i][0] = -gap_run(i) (global) or 0 (local), Y = -inf, X = H
                T V[HFORM ? 1 : R], Y[R], Hs[R];
                {
                    T gr = LOCAL ? zero
                                 : C::splat(-(a.gap_open + int32_t(i0) * a.gap_extend));
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        Hs[r] = HFORM ? gr - go : gr;
                        if constexpr (!HFORM) V[r] = gr;
                        Y[r] = neg;
                        if (!LOCAL) gr = gr - ge;
                    }
                }
                // H[i0][j - 1] (minus go in the H form): the top row's diagonal, from j = 0
                T hprev;
                if (LOCAL || i0 == 0) {
                    hprev = HFORM ? zero - go : zero;
                } else {
                    const T h0 = C::splat(-(a.gap_open + int32_t(i0 - 1) * a.gap_extend));
                    hprev = HFORM ? h0 - go : h0;
                }
                T g0 = C::splat(-a.gap_open);  // -gap_run(j + 1) on row 0
                // rows of this stripe any query of the workgroup reaches (the last stripe
                // stops after them, in steps of 4 rows)
                const uint32_t rl = __builtin_amdgcn_readfirstlane(
                    rows - i0 < uint32_t(R) ? rows - i0 : uint32_t(R));
                uint32_t word = 0;
                for (uint32_t j = 0; j < n; ++j) {
                    if ((j & 3u) == 0) word = wp[uint64_t(j >> 2) * npad];
                    const uint32_t rc = (word >> (8 * (j & 3u))) & 0xFFu;
                    const uint32_t* pcol = sprof + rc;
                    uint32_t* bj = bnd + uint64_t(j) * NB * kThreads;
                    T up, hd, xup;  // H form: Hg_up ; U/V form: U_up
                    if (i0 == 0) {
                        xup = neg;
                        if (LOCAL) {
                            up = HFORM ? zero - go : zero;
                            hd = up;
                        } else {
                            up = HFORM ? g0 - go : g0;
                            hd = hprev;
                            hprev = up;
                            g0 = g0 - ge;
                        }
                    } else {
                        const T hup = C::load(bj[0]);
                        xup = C::load(bj[kThreads]);
                        if constexpr (HFORM) up = hup;
                        else up = C::load(bj[2 * kThreads]);
                        hd = hprev;
                        hprev = hup;
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if ((r & 3) == 0 && uint32_t(r) >= rl) break;  // uniform
                        const T s = C::load(pcol[r * 32]);
                        T mv = hd + s;
                        if (LOCAL) mv = C::vmax(mv, zero);
                        if constexpr (HFORM) {
                            const T xv = C::vmax(up, xup - ge);
                            const T yv = C::vmax(Hs[r], Y[r] - ge);
                            hd = Hs[r];
                            Hs[r] = C::vmax(C::vmax(mv, xv), yv) - go;
                            up = Hs[r];
                            xup = xv;
                            Y[r] = yv;
                        } else {
                            const T xv = C::vmax(up - go, xup - ge);
                            const T yv = C::vmax(V[r] - go, Y[r] - ge);
                            hd = Hs[r];
                            up = C::vmax(mv, yv);
                            xup = xv;
                            V[r] = C::vmax(mv, xv);
                            Y[r] = yv;
                            Hs[r] = C::vmax(up, xv);
                        }
                        if (LOCAL) best = C::vmax(best, mv);
                    }
                    if (sidx + 1 < S) {  // the bottom row (rl == R) for the next stripe's top row
                        bj[0] = __builtin_bit_cast(uint32_t, Hs[R - 1]);
                        bj[kThreads] = __builtin_bit_cast(uint32_t, xup);
                        if constexpr (!HFORM) bj[2 * kThreads] = __builtin_bit_cast(uint32_t, up);
                    }
                }
                if (!LOCAL) {
#pragma unroll
                    for (int h = 0; h < Q; ++h) {
                        if (m[h] == 0 || (m[h] - 1) / R != sidx) continue;
                        const uint32_t rl = (m[h] - 1) % R;
                        int32_t v = C::half(Hs[0], h);
#pragma unroll
                        for (int r = 1; r < R; ++r)
                            if (uint32_t(r) == rl) v = C::half(Hs[r], h);
                        res[h] = HFORM ? v + a.gap_open : v;
                    }
                }
            }

#pragma unroll
            for (int h = 0; h < Q; ++h) {
                int32_t v = res[h];
                if (LOCAL) v = C::half(best, h);
                else if (m[h] == 0) v = n == 0 ? 0 : -(a.gap_open + int32_t(n - 1) * a.gap_extend);
                if (live) {
                    if (a.dist) v = int32_t(a.q_self[qi[h]] + a.r_self[k] - 2 * int64_t(v));
                    stage[h][k - c * kChunk] = v;
                }
            }
        }
        __syncthreads();

        const uint64_t col0 = uint64_t(c) * kChunk;
        const uint64_t lim64 = a.n_refs - col0;
        const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
#pragma unroll
        for (int h = 0; h < Q; ++h) {
            if (!valid[h]) continue;
            int32_t* row = a.out + uint64_t(qi[h]) * a.n_refs + col0;
            for (uint32_t t = tid; t < lim; t += kThreads) row[t] = stage[h][t];
        }
    }
}

// Exact int64 restatement of the reference DP for one pair; buf holds 6 * (min(m,n)+1).
__device__ int64_t gotoh_i64(const uint8_t* q, uint32_t m, const uint8_t* r, uint32_t n,
                             const int32_t* __restrict__ sim, const int32_t* __restrict__ simT,
                             uint32_t A, bool local, int64_t go, int64_t ge, int64_t* buf) {
    const int64_t kNeg = INT64_MIN / 4;
    if (m == 0 || n == 0) {
        if (local) return 0;
        const uint32_t L = m ? m : n;
        return L == 0 ? 0 : -(go + int64_t(L - 1) * ge);
    }
    const int32_t* S = sim;
    if (n > m) {  // keep the row buffers over the shorter sequence: s(q,r;sim) = s(r,q;sim^T)
        const uint8_t* tq = q;
        q = r;
        r = tq;
        const uint32_t tm = m;
        m = n;
        n = tm;
        S = simT;
    }
    int64_t* Mp = buf;
    int64_t* Xp = Mp + (n + 1);
    int64_t* Yp = Xp + (n + 1);
    int64_t* Mc = Yp + (n + 1);
    int64_t* Xc = Mc + (n + 1);
    int64_t* Yc = Xc + (n + 1);
    Mp[0] = 0;
    Xp[0] = kNeg;
    Yp[0] = kNeg;
    for (uint32_t j = 1; j <= n; ++j) {
        Mp[j] = local ? 0 : kNeg;
        Xp[j] = kNeg;
        Yp[j] = local ? kNeg : -(go + int64_t(j - 1) * ge);
    }
    int64_t best = local ? 0 : kNeg;
    for (uint32_t i = 1; i <= m; ++i) {
        Mc[0] = local ? 0 : kNeg;
        Xc[0] = local ? kNeg : -(go + int64_t(i - 1) * ge);
        Yc[0] = kNeg;
        const int32_t* srow = S + uint32_t(q[i - 1]) * A;
        for (uint32_t j = 1; j <= n; ++j) {
            const int64_t s = srow[r[j - 1]];
            const int64_t diag = max(Mp[j - 1], max(Xp[j - 1], Yp[j - 1]));
            int64_t mv = diag + s;
            if (local && mv < 0) mv = 0;
            Mc[j] = mv;
            Xc[j] = max(Mp[j] - go, max(Xp[j] - ge, Yp[j] - go));
            Yc[j] = max(Mc[j - 1] - go, max(Yc[j - 1] - ge, Xc[j - 1] - go));
            if (local && mv > best) best = mv;
        }
        int64_t* t;
        t = Mp; Mp = Mc; Mc = t;
        t = Xp; Xp = Xc; Xc = t;
        t = Yp; Yp = Yc; Yc = t;
    }
    if (local) return best;
    return max(Mp[n], max(Xp[n], Yp[n]));
}

__global__ __launch_bounds__(256) void gotoh_generic_kernel(GotohArgs a,
                                                            const uint32_t* __restrict__ list,
                                                            uint32_t count,
                                                            int64_t* __restrict__ scratch,
                                                            uint64_t stride) {
    const uint64_t slots = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t* buf = scratch + t * 6 * stride;
    const uint64_t total = uint64_t(count) * a.n_refs;
    for (uint64_t p = t; p < total; p += slots) {
        const uint32_t qi = list[p / a.n_refs];
        const uint64_t k = p % a.n_refs;
        const uint64_t q0 = a.q_off[qi], r0 = a.r_off[k];
        const int64_t s = gotoh_i64(a.q_codes + q0, uint32_t(a.q_off[qi + 1] - q0),
                                    a.r_codes + r0, uint32_t(a.r_off[k + 1] - r0), a.sim, a.simT,
                                    a.A, a.local != 0, a.gap_open, a.gap_extend, buf);
        const int64_t v = a.dist ? a.q_self[qi] + a.r_self[k] - 2 * s : s;
        if (v < INT32_MIN || v > INT32_MAX) *a.overflow = 1;
        a.out[uint64_t(qi) * a.n_refs + k] = int32_t(v);
    }
}

__global__ __launch_bounds__(256) void gotoh_self_kernel(GotohArgs a,
                                                         const uint8_t* __restrict__ codes,
                                                         const uint64_t* __restrict__ off,
                                                         uint64_t n, int64_t* __restrict__ res,
                                                         int64_t* __restrict__ scratch,
                                                         uint64_t stride) {
    const uint64_t slots = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t* buf = scratch + t * 6 * stride;
    for (uint64_t x = t; x < n; x += slots) {
        const uint64_t o = off[x];
        const uint32_t L = uint32_t(off[x + 1] - o);
        res[x] = gotoh_i64(codes + o, L, codes + o, L, a.sim, a.simT, a.A, a.local != 0,
                           a.gap_open, a.gap_extend, buf);
    }
}

// Self-scores s(x, x) of strings of at most kSelfMaxLen residues: one thread per string, its
// three DP rows (M, X, Y over the string, int32) in LDS, column-interleaved across the wave so
// that lanes touching the same column hit distinct banks, with the similarity table and the
// lane's own codes (also interleaved) in LDS beside them. Rows are updated in place: the cell
// (i, j) reads row i - 1 at j from LDS and keeps row i - 1 at j - 1 (the diagonal) and row i at
// j - 1 (the left neighbour) in registers. Same recurrence, boundary values and order of max as
// gotoh_i64; the host runs it only when every value stays far inside int32 (kNeg32 and a few
// gap steps below it included), so the results are equal. Dynamic LDS: A x A table, codes
// [W][64] bytes, rows [3][W][64] int32 with W = longest string + 1.
constexpr uint32_t kSelfMaxLen = 160;
constexpr uint32_t kSelfThreads = 64;

size_t self_lds_bytes(uint32_t W) {
    return kMaxAlphabet * kMaxAlphabet * sizeof(int32_t) + ((size_t(W) * kSelfThreads + 15) & ~size_t(15)) +
           3 * size_t(W) * kSelfThreads * sizeof(int32_t);
}

// Both sets in one launch: workgroups [0, nb0) take set 0, the rest set 1, so the two sets'
// strings run side by side (a set of a few thousand strings fills only a few dozen CUs).
template <bool LOCAL>
__global__ __launch_bounds__(kSelfThreads) void gotoh_self_lds_kernel(GotohArgs a, SelfSets sets,
                                                                     uint32_t W) {
    extern __shared__ int32_t self_lds[];
    constexpr int32_t kNeg32 = -(1 << 29);
    const uint32_t lane = threadIdx.x;
    const uint32_t A = a.A;
    int32_t* sim_l = self_lds;
    for (uint32_t e = lane; e < A * A; e += kSelfThreads) sim_l[e] = a.sim[e];
    __syncthreads();  // the only barrier
    uint8_t* qc = reinterpret_cast<uint8_t*>(self_lds + kMaxAlphabet * kMaxAlphabet);
    int32_t* rows = reinterpret_cast<int32_t*>(qc + ((W * kSelfThreads + 15) & ~15u));
    const bool second = blockIdx.x >= sets.nb0;
    const uint8_t* codes = second ? sets.codes[1] : sets.codes[0];
    const uint64_t* off = second ? sets.off[1] : sets.off[0];
    const uint64_t n = second ? sets.n[1] : sets.n[0];
    int64_t* res = second ? sets.res[1] : sets.res[0];
    const uint64_t x = uint64_t(second ? blockIdx.x - sets.nb0 : blockIdx.x) * kSelfThreads + lane;
    if (x >= n) return;
    const uint64_t o = off[x];
    const uint32_t L = uint32_t(off[x + 1] - o);
    const int32_t go = a.gap_open, ge = a.gap_extend;
    if (L == 0) {
        res[x] = 0;
        return;
    }
    for (uint32_t c = 0; c < L; ++c) qc[c * kSelfThreads + lane] = codes[o + c];
    auto M = [&](uint32_t j) -> int32_t& { return rows[(0 * W + j) * kSelfThreads + lane]; };
    auto X = [&](uint32_t j) -> int32_t& { return rows[(1 * W + j) * kSelfThreads + lane]; };
    auto Y = [&](uint32_t j) -> int32_t& { return rows[(2 * W + j) * kSelfThreads + lane]; };
    M(0) = 0;
    X(0) = kNeg32;
    Y(0) = kNeg32;
    for (uint32_t j = 1; j <= L; ++j) {
        M(j) = LOCAL ? 0 : kNeg32;
        X(j) = kNeg32;
        Y(j) = LOCAL ? kNeg32 : -(go + int32_t(j - 1) * ge);
    }
    int32_t best = LOCAL ? 0 : kNeg32;
    for (uint32_t i = 1; i <= L; ++i) {
        // column 0 of row i, and row i - 1 at column 0 as the first diagonal
        int32_t dm = M(0), dx = X(0), dy = Y(0);
        int32_t lm = LOCAL ? 0 : kNeg32;
        int32_t lx = LOCAL ? kNeg32 : -(go + int32_t(i - 1) * ge);
        int32_t ly = kNeg32;
        M(0) = lm;
        X(0) = lx;
        Y(0) = ly;
        const int32_t* srow = sim_l + uint32_t(qc[(i - 1) * kSelfThreads + lane]) * A;
        for (uint32_t j = 1; j <= L; ++j) {
            const int32_t um = M(j), ux = X(j), uy = Y(j);
            const int32_t sv = srow[qc[(j - 1) * kSelfThreads + lane]];
            int32_t mv = max(dm, max(dx, dy)) + sv;
            if (LOCAL && mv < 0) mv = 0;
            const int32_t xv = max(um - go, max(ux - ge, uy - go));
            const int32_t yv = max(lm - go, max(ly - ge, lx - go));
            if (LOCAL && mv > best) best = mv;
            M(j) = mv;
            X(j) = xv;
            Y(j) = yv;
            dm = um;
            dx = ux;
            dy = uy;
            lm = mv;
            lx = xv;
            ly = yv;
        }
    }
    res[x] = LOCAL ? best : max(M(L), max(X(L), Y(L)));
}

// Intra-pair wavefront kernel for long pairs that are few: one wave per (query, reference)
// pair, so 64 lanes share one DP instead of one lane taking a whole pair. Lane l owns R
// consecutive query rows of the current stripe of 64 R rows, with two registers per row (the
// H form of gotoh_reg_kernel: Hg = H - go and Y; X flows down the rows in one register) and
// the lane's query codes as byte offsets into an LDS table tab[ref code][query code] = s + go.
// The lanes form a pipeline over the reference columns: at step t lane l scores column
// j = t - l + 1 for its rows, taking the row above (Hg and X at column j) from lane l - 1's
// bottom row of step t - 1 and the diagonal Hg from the value it received one step earlier,
// both by a one-lane shuffle, and the residue r_j the same way; lane 0 takes them from the
// stripe's top row (row 0's boundary, or the previous stripe's bottom row kept by lane 63 in
// the wave's boundary buffer) and from the reference, read in 64-column blocks every 64 steps.
// A stripe takes n + 63 steps and has no barrier; between stripes one workgroup-scope
// release / acquire fence orders lane 63's boundary stores before the next stripe's loads.
// Rows past the query's end compute values nobody reads: they are masked out of the local
// maximum, and the global score is read at row m, column n. int32 cells (host range bound:
// (q_max + r_max + 2) * max(|s|, go, ge) < 2^28) with kWaveNeg as -inf, so every comparison is
// the int64 DP's. Same recurrences, boundary values and results as gotoh_i64 when
// gap_open >= gap_extend.
constexpr int32_t kWaveNeg = -(1 << 29);
constexpr uint32_t kWaveThreads = 256;  // four independent waves per workgroup

template <int R, bool LOCAL>
__global__ __launch_bounds__(kWaveThreads) void gotoh_wave_kernel(GotohArgs a, GotohWaveJob job,
                                                                  int32_t* __restrict__ bnd_all,
                                                                  uint32_t W) {
    __shared__ int32_t tab[kMaxAlphabet * kMaxAlphabet];  // [ref code][query code] = s + go
    const int32_t go = a.gap_open, ge = a.gap_extend;
    const uint32_t A = a.A;
    for (uint32_t e = threadIdx.x; e < kMaxAlphabet * kMaxAlphabet; e += kWaveThreads) {
        const uint32_t rc = e / kMaxAlphabet, qc = e % kMaxAlphabet;
        tab[e] = (rc < A && qc < A) ? a.sim[qc * A + rc] + go : 0;
    }
    __syncthreads();  // the only barrier: the waves below are independent
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t gw = uint64_t(blockIdx.x) * (kWaveThreads / 64) + (threadIdx.x >> 6);
    const uint64_t nw = uint64_t(gridDim.x) * (kWaveThreads / 64);
    int32_t* bH = bnd_all ? bnd_all + gw * 2 * (uint64_t(W) + 1) : nullptr;
    int32_t* bX = bH ? bH + (uint64_t(W) + 1) : nullptr;
    constexpr uint32_t S = 64 * R;  // rows per stripe

    for (uint64_t p = gw; p < job.n_pairs; p += nw) {
        uint32_t qi, ri;
        const uint8_t *q, *r;
        uint32_t m, n;
        if (job.list) {  // matrix: every listed query against every reference
            qi = job.list[p / a.n_refs];
            ri = uint32_t(p % a.n_refs);
            const uint64_t q0 = a.q_off[qi], r0 = a.r_off[ri];
            q = a.q_codes + q0;
            r = a.r_codes + r0;
            m = uint32_t(a.q_off[qi + 1] - q0);
            n = uint32_t(a.r_off[ri + 1] - r0);
        } else {  // self scores of two sets
            const int set = p < job.n_first ? 0 : 1;
            qi = ri = uint32_t(set ? p - job.n_first : p);
            const uint64_t o = job.off[set][qi];
            q = r = job.codes[set] + o;
            m = n = uint32_t(job.off[set][qi + 1] - o);
        }
        if (m > job.q_max || n > job.r_max) continue;  // outside the bounds the host sized for
        int64_t score;
        if (m == 0 || n == 0) {
            const uint32_t L = m ? m : n;
            score = LOCAL || L == 0 ? 0 : -(int64_t(go) + int64_t(L - 1) * ge);
        } else {
            const uint32_t ns = (m + S - 1) / S;
            int32_t best = 0, res = 0;
            for (uint32_t st = 0; st < ns; ++st) {
                const uint32_t row0 = st * S + lane * R;  // the lane's rows: row0 + 1 ..
                uint32_t qo[R];
                int32_t Hs[R], Y[R];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const uint32_t i = row0 + k + 1;  // 1-based query row
                    qo[k] = i <= m ? uint32_t(q[i - 1]) : 0u;
                    // column 0: H = X = -gap_run(i) (global) or 0 (local); Y unreachable
                    const int32_t h0 = LOCAL ? 0 : -(go + int32_t(i - 1) * ge);
                    Hs[k] = h0 - go;
                    Y[k] = kWaveNeg;
                }
                // the row above the stripe at column 0 (lane 0's first diagonal)
                int32_t top0;
                if (st == 0) top0 = -go;  // H[0][0] = 0
                else top0 = bH[0];
                // this lane's bottom row at column 0, until the lane starts its columns
                int32_t outH = Hs[R - 1], outX = kWaveNeg;
                if (st + 1 < ns) {
                    // top0's load completes before lane 63 overwrites bH[0] for the next stripe
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
                    if (lane == 63) bH[0] = outH;
                }
                int32_t prevIn = lane == 0 ? top0 : 0;
                uint32_t rcur = 0;
                int32_t rblk = 0, hblk = 0, xblk = 0;
                const uint32_t steps = n + 63;
                for (uint32_t t = 0; t < steps; ++t) {
                    if ((t & 63u) == 0) {  // lane 0's sources for the next 64 columns
                        const uint32_t c = t + lane;  // 0-based column index j - 1
                        rblk = c < n ? int32_t(r[c]) : 0;
                        if (st > 0 && c < n) {
                            hblk = bH[c + 1];
                            xblk = bX[c + 1];
                        }
                    }
                    const int32_t jt = int32_t(t) + 1;  // lane 0's column this step
                    int32_t topH, topX;
                    if (st == 0) {
                        topH = (LOCAL ? 0 : -(go + (jt - 1) * ge)) - go;
                        topX = kWaveNeg;
                    } else {
                        topH = __builtin_amdgcn_readlane(hblk, int(t & 63u));
                        topX = __builtin_amdgcn_readlane(xblk, int(t & 63u));
                    }
                    const uint32_t r0c = uint32_t(__builtin_amdgcn_readlane(rblk, int(t & 63u)));
                    int32_t inH = __shfl_up(outH, 1, 64), inX = __shfl_up(outX, 1, 64);
                    uint32_t rc = uint32_t(__shfl_up(int32_t(rcur), 1, 64));
                    if (lane == 0) {
                        inH = topH;
                        inX = topX;
                        rc = r0c;
                    }
                    rcur = rc;
                    const int32_t j = int32_t(t) - int32_t(lane) + 1;
                    if (j >= 1 && j <= int32_t(n)) {
                        const int32_t* trow = tab + rc * kMaxAlphabet;
                        int32_t up = inH, xup = inX, hd = prevIn;
#pragma unroll
                        for (int k = 0; k < R; ++k) {
                            int32_t mv = hd + trow[qo[k]];
                            if (LOCAL) mv = max(mv, 0);
                            const int32_t xv = max(up, xup - ge);
                            const int32_t yv = max(Hs[k], Y[k] - ge);
                            hd = Hs[k];
                            Hs[k] = max(max(mv, xv), yv) - go;
                            up = Hs[k];
                            xup = xv;
                            Y[k] = yv;
                            if (LOCAL) {
                                if (row0 + k < m) best = max(best, mv);
                            } else if (row0 + k + 1 == m && j == int32_t(n)) {
                                res = Hs[k] + go;
                            }
                        }
                        outH = up;
                        outX = xup;
                        if (lane == 63 && st + 1 < ns) {
                            bH[j] = outH;
                            bX[j] = outX;
                        }
                    }
                    prevIn = inH;
                }
                if (st + 1 < ns) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
            }
            if (LOCAL) {
                for (int off = 32; off > 0; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
                score = best;
            } else {
                const uint32_t lane_m = ((m - 1) % S) / R;
                score = __shfl(res, int(lane_m), 64);
            }
        }
        if (lane == 0) {
            if (!job.list) {
                job.res[p < job.n_first ? 0 : 1][qi] = score;
            } else {
                const int64_t v = a.dist ? a.q_self[qi] + a.r_self[ri] - 2 * score : score;
                if (v < INT32_MIN || v > INT32_MAX) *a.overflow = 1;
                a.out[uint64_t(qi) * a.n_refs + ri] = int32_t(v);
            }
        }
    }
}

template <int R>
void launch_wave_rows(const GotohArgs& a, const GotohWaveJob& job, uint32_t grid, int32_t* bnd,
                      uint32_t W, hipStream_t s) {
    if (a.local)
        hipLaunchKernelGGL((gotoh_wave_kernel<R, true>), dim3(grid), dim3(kWaveThreads), 0, s, a,
                           job, bnd, W);
    else
        hipLaunchKernelGGL((gotoh_wave_kernel<R, false>), dim3(grid), dim3(kWaveThreads), 0, s,
                           a, job, bnd, W);
}

template <class C, int R, bool HFORM>
void launch_form(const GotohArgs& a, const uint32_t* list, uint32_t count, hipStream_t s) {
    const uint64_t nch = a.refs.nchunks;
    const uint64_t groups = (uint64_t(count) + C::Q - 1) / C::Q;
    const uint64_t max_g = (uint64_t(1) << 30) / nch;
    for (uint64_t g0 = 0; g0 < groups; g0 += max_g) {
        const uint64_t ng = groups - g0 < max_g ? groups - g0 : max_g;
        if (a.local)
            hipLaunchKernelGGL((gotoh_reg_kernel<C, R, true, HFORM>), dim3(uint32_t(ng * nch)),
                               dim3(kThreads), 0, s, a, list, count, g0);
        else
            hipLaunchKernelGGL((gotoh_reg_kernel<C, R, false, HFORM>), dim3(uint32_t(ng * nch)),
                               dim3(kThreads), 0, s, a, list, count, g0);
    }
}

// The packed int16 cell always has gap_open >= gap_extend (the host routes the other case
// to int32), so only the int32 cell instantiates the U/V form.
template <class C, int R>
void launch_rows(const GotohArgs& a, const uint32_t* list, uint32_t count, hipStream_t s) {
    if constexpr (C::Q == 1) {
        if (a.gap_open < a.gap_extend) {
            launch_form<C, R, false>(a, list, count, s);
            return;
        }
    }
    launch_form<C, R, true>(a, list, count, s);
}

template <class C>
void launch_cell(const GotohArgs& a, int rows, const uint32_t* list, uint32_t count,
                 hipStream_t s) {
    switch (rows) {
        case 4: launch_rows<C, 4>(a, list, count, s); break;
        case 8: launch_rows<C, 8>(a, list, count, s); break;
        case 12: launch_rows<C, 12>(a, list, count, s); break;
        case 16: launch_rows<C, 16>(a, list, count, s); break;
        case 20: launch_rows<C, 20>(a, list, count, s); break;
        case 24: launch_rows<C, 24>(a, list, count, s); break;
        case 28: launch_rows<C, 28>(a, list, count, s); break;
        case 32: launch_rows<C, 32>(a, list, count, s); break;
        case 36: launch_rows<C, 36>(a, list, count, s); break;
        case 40: launch_rows<C, 40>(a, list, count, s); break;
        case 44: launch_rows<C, 44>(a, list, count, s); break;
        case 48: launch_rows<C, 48>(a, list, count, s); break;
        case 52: launch_rows<C, 52>(a, list, count, s); break;
        case 56: launch_rows<C, 56>(a, list, count, s); break;
        case 60: launch_rows<C, 60>(a, list, count, s); break;
        default: launch_rows<C, 64>(a, list, count, s); break;
    }
}

template <class C, bool HFORM>
void launch_stripes_form(const GotohArgs& a, const uint32_t* list, uint32_t count,
                         uint64_t n_units, uint32_t grid, uint32_t* bnd, uint32_t W,
                         uint32_t q_max, hipStream_t s) {
    // the whole profile when it fits kStripeFullRows rows, else one stripe at a time
    const uint32_t q_pad = (q_max + kStripeRows - 1) / kStripeRows * kStripeRows;
    const uint32_t rows = q_pad <= kStripeFullRows ? q_pad : uint32_t(kStripeRows);
    const size_t lds = size_t(rows) * 32 * sizeof(uint32_t);
    if (a.local)
        hipLaunchKernelGGL((gotoh_stripes_kernel<C, kStripeRows, true, HFORM>), dim3(grid),
                           dim3(kThreads), lds, s, a, list, count, n_units, bnd, W, rows);
    else
        hipLaunchKernelGGL((gotoh_stripes_kernel<C, kStripeRows, false, HFORM>), dim3(grid),
                           dim3(kThreads), lds, s, a, list, count, n_units, bnd, W, rows);
}

}  // namespace

uint64_t gotoh_stripes_units(uint32_t count, bool packed16, uint32_t nchunks) {
    const uint64_t q = packed16 ? 2 : 1;
    return (uint64_t(count) + q - 1) / q * nchunks;
}

uint64_t gotoh_stripes_words(uint32_t grid, uint32_t W, bool hform) {
    return uint64_t(grid) * (hform ? 2 : 3) * W * kThreads;
}

void launch_gotoh_stripes(const GotohArgs& a, bool packed16, const uint32_t* list,
                          uint32_t count, uint32_t grid, uint32_t* bnd, uint32_t W,
                          uint32_t q_max, hipStream_t s) {
    if (count == 0 || a.refs.nchunks == 0 || grid == 0) return;
    const uint64_t units = gotoh_stripes_units(count, packed16, a.refs.nchunks);
    if (packed16)
        launch_stripes_form<I16x2, true>(a, list, count, units, grid, bnd, W, q_max, s);
    else if (a.gap_open < a.gap_extend)
        launch_stripes_form<I32, false>(a, list, count, units, grid, bnd, W, q_max, s);
    else launch_stripes_form<I32, true>(a, list, count, units, grid, bnd, W, q_max, s);
}

void launch_gotoh_fast(const GotohArgs& a, int rows, bool packed16, const uint32_t* list,
                       uint32_t count, hipStream_t s) {
    if (count == 0 || a.refs.nchunks == 0) return;
    if (packed16) launch_cell<I16x2>(a, rows, list, count, s);
    else launch_cell<I32>(a, rows, list, count, s);
}

void launch_gotoh_generic(const GotohArgs& a, const uint32_t* list, uint32_t count,
                          int64_t* scratch, uint32_t slots, uint64_t stride, hipStream_t s) {
    if (count == 0 || a.n_refs == 0) return;
    hipLaunchKernelGGL(gotoh_generic_kernel, dim3(slots / 256), dim3(256), 0, s, a, list, count,
                       scratch, stride);
}

void launch_gotoh_self(const GotohArgs& a, const uint8_t* codes, const uint64_t* off, uint64_t n,
                       int64_t* res, int64_t* scratch, uint32_t slots, uint64_t stride,
                       hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(gotoh_self_kernel, dim3(slots / 256), dim3(256), 0, s, a, codes, off, n,
                       res, scratch, stride);
}

uint32_t gotoh_wave_rows(uint32_t q_max) {
    return q_max > 1024 ? 32u : q_max > 512 ? 16u : 8u;
}

uint64_t gotoh_wave_bnd_words(uint32_t grid, uint32_t q_max, uint32_t r_max) {
    // one stripe needs no boundary buffer
    if (q_max <= 64 * gotoh_wave_rows(q_max)) return 0;
    return uint64_t(grid) * (kWaveThreads / 64) * 2 * (uint64_t(r_max) + 1);
}

void launch_gotoh_wave(const GotohArgs& a, const GotohWaveJob& job, uint32_t q_max,
                       uint32_t r_max, uint32_t grid, int32_t* bnd, hipStream_t s) {
    if (job.n_pairs == 0 || grid == 0) return;
    switch (gotoh_wave_rows(q_max)) {
        case 8: launch_wave_rows<8>(a, job, grid, bnd, r_max, s); break;
        case 16: launch_wave_rows<16>(a, job, grid, bnd, r_max, s); break;
        default: launch_wave_rows<32>(a, job, grid, bnd, r_max, s); break;
    }
}

bool gotoh_self_fits_lds(uint32_t max_len, int64_t mx) {
    // every finite value is within (2 max_len + 2) * mx of 0; sentinel-derived values stay
    // within kNeg32 - (2 max_len + 2) * mx; both far inside int32 while that span is < 2^28
    return max_len <= kSelfMaxLen && (2 * int64_t(max_len) + 2) * mx < (int64_t(1) << 28);
}

void launch_gotoh_self_lds(const GotohArgs& a, const SelfSets& sets, uint32_t max_len,
                           hipStream_t s) {
    SelfSets t = sets;
    t.nb0 = uint32_t((sets.n[0] + kSelfThreads - 1) / kSelfThreads);
    const uint32_t nb1 = uint32_t((sets.n[1] + kSelfThreads - 1) / kSelfThreads);
    if (t.nb0 + nb1 == 0) return;
    const uint32_t W = max_len + 1;
    const size_t lds = self_lds_bytes(W);
    if (a.local)
        hipLaunchKernelGGL(gotoh_self_lds_kernel<true>, dim3(t.nb0 + nb1), dim3(kSelfThreads),
                           lds, s, a, t, W);
    else
        hipLaunchKernelGGL(gotoh_self_lds_kernel<false>, dim3(t.nb0 + nb1), dim3(kSelfThreads),
                           lds, s, a, t, W);
}

}  // namespace st
