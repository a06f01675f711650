// Single-pair kernels behind the scalar entry points: align_score (reference
// src/pairwise.cpp:131-139), align_pair (src/pairwise.cpp:141-267), Index::align
// (src/index.cpp:105-179), hamming and levenshtein (src/distance.cpp:68-95).
//
// Through the batch machinery a call with one pair pays for uploads through the staging
// slots, reference packing, several launches and downloads (70-220 us). Here the pair travels
// in the kernel arguments, one wave computes it and the results land in pinned host memory, so
// a call is one launch and one stream synchronisation.
//
// Gotoh (scores and tracebacks) runs the DP as a wavefront over the 64 lanes: lane l owns query
// row l + 1 (queries of at most kSingleMaxQ residues) and at step t scores column
// j = t - l + 1, taking the cell above from lane l - 1's cell of the previous step by a one-lane
// shuffle and the diagonal from what it received one step earlier. The recurrences, boundary
// values, argmax order (M, then X, then Y, each replacing only on a strictly larger value) and
// back-pointer bytes are traceback_kernel's, with int32 cells under the host's range bound
// ((m + n + 2) * max(|s|, go, ge) < 2^28, sentinel -2^29), so every comparison and tie is the
// int64 DP's. The local start is the first maximal M in row-major order: each lane keeps its
// row's first maximum (strict >), and the smallest row holding the wave's maximum wins. The
// traceback's direction bytes sit in LDS, [m + 1][n + 1] row-major, and lane 0 walks them with
// the reference's rules. The penalty form of Index::align is the same kernel in global mode on
// the negated penalty table (the host negates; the score is then minus the penalty).
// Levenshtein is the same wavefront with unit costs; Hamming one ballot per 64 positions.

#include "st_internal.hpp"

namespace st {
namespace {

constexpr uint32_t kSingleLanes = 64;
constexpr int32_t kSingleNeg = -(1 << 29);
constexpr uint8_t kSM = 0, kSX = 1, kSY = 2, kSStop = 3;
constexpr uint8_t kSZero = 0x40;

__device__ __forceinline__ int32_t pick3(int32_t vm, int32_t vx, int32_t vy, uint8_t& who) {
    int32_t v = vm;
    who = kSM;
    if (vx > v) {
        v = vx;
        who = kSX;
    }
    if (vy > v) {
        v = vy;
        who = kSY;
    }
    return v;
}

__device__ __forceinline__ int32_t run_cost(uint32_t L, int32_t go, int32_t ge) {
    return L == 0 ? 0 : go + int32_t(L - 1) * ge;
}

template <bool LOCAL, bool TRACE>
__global__ __launch_bounds__(kSingleLanes) void single_gotoh_kernel(SingleArgs a) {
    __shared__ int32_t sim_l[kMaxAlphabet * kMaxAlphabet];
    __shared__ uint8_t q_l[kSingleMaxQ], r_l[kSingleMaxR];
    __shared__ uint8_t dirs[TRACE ? (kSingleMaxQ + 1) * (kSingleMaxR + 1) : 1];
    __shared__ uint8_t ops_l[TRACE ? kSingleMaxQ + kSingleMaxR : 1];
    const uint32_t lane = threadIdx.x;
    const uint32_t m = a.m, n = a.n, A = a.A, W = n + 1;
    const int32_t go = a.go, ge = a.ge;
    for (uint32_t e = lane; e < A * A; e += kSingleLanes) sim_l[e] = a.sim[e];
    for (uint32_t e = lane; e < m; e += kSingleLanes) q_l[e] = a.q[e];
    for (uint32_t e = lane; e < n; e += kSingleLanes) r_l[e] = a.r[e];
    if (TRACE) {  // row 0 and column 0: the boundary back-pointers of traceback_kernel
        for (uint32_t c = lane; c <= n; c += kSingleLanes)
            dirs[c] = uint8_t(kSStop | (kSM << 2) | ((c > 1 ? kSY : kSM) << 4));
        for (uint32_t row = lane + 1; row <= m; row += kSingleLanes)
            dirs[row * W] = uint8_t(kSStop | ((row > 1 ? kSX : kSM) << 2) | (kSM << 4));
    }
    __syncthreads();

    const uint32_t i = lane;  // this lane's row is i + 1
    const bool has_row = i < m;
    const int32_t* srow = sim_l + (has_row ? uint32_t(q_l[i]) : 0u) * A;
    // the lane's last cell (row i + 1, column j - 1): column 0 before the first step
    int32_t lm = LOCAL ? 0 : kSingleNeg;
    int32_t lx = LOCAL ? kSingleNeg : -run_cost(i + 1, go, ge);
    int32_t ly = kSingleNeg;
    // the diagonal (row i, column j - 1): column 0 of row i before the first step
    int32_t dm = i == 0 || LOCAL ? 0 : kSingleNeg;
    int32_t dx = i == 0 || LOCAL ? kSingleNeg : -run_cost(i, go, ge);
    int32_t dy = kSingleNeg;
    int32_t best = LOCAL ? 0 : kSingleNeg;
    uint32_t bj = 0;
    const uint32_t steps = m + n - 1;
    for (uint32_t t = 0; t < steps; ++t) {
        // the cell above (row i, column j): lane i - 1's cell of the previous step, or row 0
        int32_t um = __shfl_up(lm, 1, kSingleLanes);
        int32_t ux = __shfl_up(lx, 1, kSingleLanes);
        int32_t uy = __shfl_up(ly, 1, kSingleLanes);
        const int32_t j = int32_t(t) - int32_t(i) + 1;
        if (!has_row || j < 1 || j > int32_t(n)) continue;
        if (i == 0) {
            um = LOCAL ? 0 : kSingleNeg;
            ux = kSingleNeg;
            uy = LOCAL ? kSingleNeg : -run_cost(uint32_t(j), go, ge);
        }
        uint8_t wm, wx, wy;
        int32_t mv = pick3(dm, dx, dy, wm) + srow[r_l[j - 1]];
        if (LOCAL && mv < 0) {
            mv = 0;
            wm = kSStop;
        }
        const int32_t xv = pick3(um - go, ux - ge, uy - go, wx);
        const int32_t yv = pick3(lm - go, lx - go, ly - ge, wy);
        if (LOCAL && mv > best) {
            best = mv;
            bj = uint32_t(j);
        }
        if (TRACE)
            dirs[(i + 1) * W + uint32_t(j)] =
                uint8_t(wm | (wx << 2) | (wy << 4) | (mv == 0 ? kSZero : 0));
        lm = mv;
        lx = xv;
        ly = yv;
        dm = um;
        dx = ux;
        dy = uy;
    }

    int32_t score;
    uint8_t st = kSM;
    uint32_t ei, ej;
    bool run = true;
    if (LOCAL) {
        int32_t wbest = has_row ? best : 0;
        for (int off = 32; off > 0; off >>= 1) wbest = max(wbest, __shfl_xor(wbest, off, 64));
        const uint64_t holders = __ballot(has_row && best == wbest);
        const int first = holders ? __ffsll(static_cast<unsigned long long>(holders)) - 1 : 0;
        const uint32_t first_j = uint32_t(__shfl(int32_t(bj), first, kSingleLanes));
        score = wbest;
        run = wbest != 0;
        ei = run ? uint32_t(first) + 1 : 0;
        ej = run ? first_j : 0;
    } else {
        // the end cell (m, n) is lane m - 1's last cell
        const int32_t fm = __shfl(lm, int(m - 1), kSingleLanes);
        const int32_t fx = __shfl(lx, int(m - 1), kSingleLanes);
        const int32_t fy = __shfl(ly, int(m - 1), kSingleLanes);
        score = pick3(fm, fx, fy, st);
        ei = m;
        ej = n;
    }
    if (!TRACE) {
        if (lane == 0) a.out->score = score;
        return;
    }
    __syncthreads();  // every direction byte is in LDS
    if (lane != 0) return;
    uint32_t cnt = 0, ci = ei, cj = ej;
    while (run && (ci > 0 || cj > 0)) {
        const uint8_t dv = dirs[ci * W + cj];
        if (st == kSM) {
            if (ci == 0 || cj == 0) break;
            if (LOCAL && (dv & kSZero)) break;  // the score resets here
            ops_l[cnt++] = q_l[ci - 1] == r_l[cj - 1] ? 'M' : 'S';
            st = dv & 3u;
            --ci;
            --cj;
            if (st == kSStop) break;
        } else if (st == kSX) {
            if (ci == 0) break;
            ops_l[cnt++] = 'D';
            st = (dv >> 2) & 3u;
            --ci;
        } else {
            if (cj == 0) break;
            ops_l[cnt++] = 'I';
            st = (dv >> 4) & 3u;
            --cj;
        }
    }
    for (uint32_t t = 0; t < cnt; ++t) a.out->ops[t] = ops_l[cnt - 1 - t];
    a.out->n_ops = cnt;
    a.out->q_begin = ci;
    a.out->r_begin = cj;
    a.out->score = score;
}

// Unit-cost edit distance, rows over q (1 <= m <= kSingleMaxQ), columns over r (n >= 1).
__global__ __launch_bounds__(kSingleLanes) void single_levenshtein_kernel(SingleArgs a) {
    __shared__ uint8_t q_l[kSingleMaxQ], r_l[kSingleMaxR];
    const uint32_t lane = threadIdx.x;
    const uint32_t m = a.m, n = a.n;
    for (uint32_t e = lane; e < m; e += kSingleLanes) q_l[e] = a.q[e];
    for (uint32_t e = lane; e < n; e += kSingleLanes) r_l[e] = a.r[e];
    __syncthreads();
    const uint32_t i = lane;
    const bool has_row = i < m;
    int32_t left = int32_t(i) + 1;  // column 0 of row i + 1
    int32_t diag = int32_t(i);      // column 0 of row i
    for (uint32_t t = 0; t < m + n - 1; ++t) {
        int32_t up = __shfl_up(left, 1, kSingleLanes);
        const int32_t j = int32_t(t) - int32_t(i) + 1;
        if (!has_row || j < 1 || j > int32_t(n)) continue;
        if (i == 0) up = j;
        const int32_t cur =
            min(min(up + 1, left + 1), diag + (q_l[i] != r_l[j - 1] ? 1 : 0));
        diag = up;
        left = cur;
    }
    const int32_t d = __shfl(left, int(m - 1), kSingleLanes);
    if (lane == 0) a.out->score = d;
}

// Mismatched positions of two strings of length n (<= kSingleMaxR).
__global__ __launch_bounds__(kSingleLanes) void single_hamming_kernel(SingleArgs a) {
    const uint32_t lane = threadIdx.x;
    uint32_t count = 0;
    for (uint32_t base = 0; base < a.n; base += kSingleLanes) {
        const uint32_t k = base + lane;
        count += uint32_t(__popcll(__ballot(k < a.n && a.q[k] != a.r[k])));
    }
    if (lane == 0) a.out->score = count;
}

}  // namespace

void launch_single(SingleKind kind, const SingleArgs& a, hipStream_t s) {
    switch (kind) {
        case SingleKind::GotohGlobal:
            hipLaunchKernelGGL((single_gotoh_kernel<false, false>), dim3(1), dim3(kSingleLanes), 0,
                               s, a);
            break;
        case SingleKind::GotohLocal:
            hipLaunchKernelGGL((single_gotoh_kernel<true, false>), dim3(1), dim3(kSingleLanes), 0,
                               s, a);
            break;
        case SingleKind::TracebackGlobal:
            hipLaunchKernelGGL((single_gotoh_kernel<false, true>), dim3(1), dim3(kSingleLanes), 0,
                               s, a);
            break;
        case SingleKind::TracebackLocal:
            hipLaunchKernelGGL((single_gotoh_kernel<true, true>), dim3(1), dim3(kSingleLanes), 0,
                               s, a);
            break;
        case SingleKind::Levenshtein:
            hipLaunchKernelGGL(single_levenshtein_kernel, dim3(1), dim3(kSingleLanes), 0, s, a);
            break;
        case SingleKind::Hamming:
            hipLaunchKernelGGL(single_hamming_kernel, dim3(1), dim3(kSingleLanes), 0, s, a);
            break;
    }
}

}  // namespace st
