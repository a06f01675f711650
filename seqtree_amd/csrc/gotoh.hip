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
// Fast kernel: workgroup = (query, reference chunk); each lane owns one reference and
// sweeps its columns j, keeping a stripe of R query rows in registers (R >= m, rows past m
// are dummies whose profile entries can never win). With U = max(M, Y) and V = max(M, X)
// the cell update is 10 VALU ops (11 local):
//   M = Hdiag + s;  X = max(U_up - go, X_up - ge);  Y = max(V_left - go, Y_left - ge)
//   U = max(M, Y);  V = max(M, X);  H = max(U, X)
// The substitution score comes from a query profile in LDS, prof[row][code] with a fixed
// 32-entry row stride, so the lookup is one ds_read with an immediate row offset.
// It runs in int32 when a host-side range proof shows every reachable value, the -2^30
// sentinel and one gap step below it stay inside int32; otherwise the int64 generic kernel
// (a straight two-row DP per thread) computes exactly what the reference's int64 DP does,
// including the narrow() overflow check.

#include "st_internal.hpp"

namespace st {
namespace {

constexpr int32_t kNeg32 = -(1 << 30);

template <int R, bool LOCAL>
__global__ __launch_bounds__(kThreads) void gotoh_fast_kernel(GotohArgs a,
                                                              const uint32_t* __restrict__ list,
                                                              uint64_t base) {
    __shared__ int32_t prof[R * 32];
    __shared__ int32_t stage[kChunk];

    const uint32_t tid = threadIdx.x;
    const uint32_t nch = a.refs.nchunks;
    const uint32_t c = blockIdx.x % nch;
    const uint32_t qi = list[base + blockIdx.x / nch];
    const uint64_t q0 = a.q_off[qi];
    const uint32_t m = uint32_t(a.q_off[qi + 1] - q0);
    const uint8_t* q = a.q_codes + q0;
    const uint32_t A = a.A;

    for (uint32_t e = tid; e < uint32_t(R) * 32; e += kThreads) {
        const uint32_t r = e >> 5, b = e & 31u;
        prof[e] = (r < m && b < A) ? a.sim[uint32_t(q[r]) * A + b] : (LOCAL ? kNeg32 : 0);
    }
    __syncthreads();

    const int32_t go = a.gap_open, ge = a.gap_extend;
    const uint32_t npad = a.refs.npad;
    const int64_t qself = a.dist ? a.q_self[qi] : 0;

    for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
        const uint32_t slot = c * kChunk + it * kThreads + tid;
        const uint32_t k = a.refs.orig[slot];
        if (k == kNoSlot) continue;
        const uint32_t n = a.refs.len[slot];

        int32_t V[R], Y[R], H[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int32_t g = LOCAL ? 0 : -(go + r * ge);  // -gap_run(r + 1)
            V[r] = g;
            Y[r] = kNeg32;
            H[r] = g;
        }
        int32_t hrow0 = 0;
        int32_t best = 0;
        const uint32_t* wp = a.refs.words + slot;
        uint32_t word = 0;
        for (uint32_t j = 0; j < n; ++j) {
            if ((j & 3u) == 0) word = wp[uint64_t(j >> 2) * npad];
            const uint32_t rc = (word >> (8 * (j & 3u))) & 0xFFu;
            const int32_t* pcol = prof + rc;
            int32_t uup, hd;
            int32_t xup = kNeg32;
            if (LOCAL) {
                uup = 0;
                hd = 0;
            } else {
                const int32_t g = -(go + int32_t(j) * ge);  // row 0 of column j+1
                uup = g;
                hd = hrow0;
                hrow0 = g;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int32_t s = pcol[r * 32];
                int32_t mv = hd + s;
                if (LOCAL) mv = max(mv, 0);
                const int32_t xv = max(uup - go, xup - ge);
                const int32_t yv = max(V[r] - go, Y[r] - ge);
                hd = H[r];
                uup = max(mv, yv);
                xup = xv;
                V[r] = max(mv, xv);
                Y[r] = yv;
                H[r] = max(uup, xv);
                if (LOCAL) best = max(best, mv);
            }
        }

        int32_t res;
        if (LOCAL) {
            res = best;
        } else if (m == 0) {
            res = n == 0 ? 0 : -(go + int32_t(n - 1) * ge);
        } else {
            res = H[0];
#pragma unroll
            for (int r = 1; r < R; ++r)
                if (uint32_t(r) == m - 1) res = H[r];
        }
        if (a.dist) res = int32_t(qself + a.r_self[k] - 2 * int64_t(res));
        stage[k - c * kChunk] = res;
    }
    __syncthreads();

    const uint64_t col0 = uint64_t(c) * kChunk;
    const uint64_t lim64 = a.n_refs - col0;
    const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
    int32_t* row = a.out + uint64_t(qi) * a.n_refs + col0;
    for (uint32_t t = tid; t < lim; t += kThreads) row[t] = stage[t];
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

template <int R>
void launch_rows(const GotohArgs& a, const uint32_t* list, uint32_t count, hipStream_t s) {
    const uint64_t nch = a.refs.nchunks;
    const uint64_t max_q = (uint64_t(1) << 30) / nch;
    for (uint64_t q0 = 0; q0 < count; q0 += max_q) {
        const uint64_t nq = count - q0 < max_q ? count - q0 : max_q;
        if (a.local)
            hipLaunchKernelGGL((gotoh_fast_kernel<R, true>), dim3(uint32_t(nq * nch)),
                               dim3(kThreads), 0, s, a, list, q0);
        else
            hipLaunchKernelGGL((gotoh_fast_kernel<R, false>), dim3(uint32_t(nq * nch)),
                               dim3(kThreads), 0, s, a, list, q0);
    }
}

}  // namespace

void launch_gotoh_fast(const GotohArgs& a, int rows, const uint32_t* list, uint32_t count,
                       hipStream_t s) {
    if (count == 0 || a.refs.nchunks == 0) return;
    switch (rows) {
        case 8: launch_rows<8>(a, list, count, s); break;
        case 16: launch_rows<16>(a, list, count, s); break;
        case 24: launch_rows<24>(a, list, count, s); break;
        case 32: launch_rows<32>(a, list, count, s); break;
        case 40: launch_rows<40>(a, list, count, s); break;
        case 48: launch_rows<48>(a, list, count, s); break;
        case 56: launch_rows<56>(a, list, count, s); break;
        default: launch_rows<64>(a, list, count, s); break;
    }
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

}  // namespace st
