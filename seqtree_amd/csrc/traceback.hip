// Gotoh alignment with traceback for a list of (query, reference) pairs (reference:
// src/pairwise.cpp:141-267, `align_pair`).
//
// One lane per pair, 64 lanes per workgroup, a grid-stride loop over the pairs. The DP runs
// row by row like the reference's, with the three state rows updated in place (the cell
// (i, j) reads row i - 1 at j, keeps row i - 1 at j - 1 and row i at j - 1 in registers) and
// the same argmax order (M, then X, then Y, each replacing only on a strictly larger value).
// Every cell's back-pointers go to a direction byte: bits 0-1 the state M came from (or stop),
// 2-3 X's, 4-5 Y's, bit 6 set when M is 0 (the local traceback stops there). Rows and
// direction bytes are lane-interleaved ([.. ][64 lanes]), so lanes at the same cell touch
// consecutive addresses. The traceback then follows the reference's rules from the end cell
// and writes the ops ('M' equal codes, 'S', 'D' query residue against a gap, 'I' reference
// residue against a gap) in forward order.
//
// T is int32 when the host's range bound holds (every value far inside int32, the sentinel
// -2^29 and a few gap steps below it included, so every comparison and tie is the int64
// one's) and int64 otherwise, with the reference's own -inf = INT64_MIN / 4.
//
// The same kernels run the penalty form of src/index.cpp:105-179 (`Index::align`): a global
// alignment minimising substitution penalty plus affine gap cost, with its minimum over
// (M, X, Y) keeping the first of equal values. Negating the penalties turns that minimum into
// this maximum with the same state order and the same strict comparisons, and the boundary
// back-pointers are the same, so the host passes sim = -penalty, and the kernels write -score.
// The int64 cells then use minus the reference's int32 kInf (INT32_MAX / 4) as -inf, which
// makes every int64 sum the reference's int32 one wherever the latter does not overflow.
//
// A pair longer than the host's bounds (q_max, r_max) would overrun the per-workgroup
// direction scratch; it is skipped and raises the too_long flag, which the host turns into an
// error.

#include "st_internal.hpp"

namespace st {
namespace {

constexpr uint32_t kTbLanes = 64;
constexpr uint8_t kStM = 0, kStX = 1, kStY = 2, kStStop = 3;
constexpr uint8_t kZeroBit = 0x40;

template <class T>
struct NegInf;
template <>
struct NegInf<int32_t> {
    static constexpr int32_t v = -(1 << 29);
};
template <>
struct NegInf<int64_t> {
    static constexpr int64_t v = INT64_MIN / 4;
};

template <class T>
__device__ __forceinline__ T pick(T vm, T vx, T vy, uint8_t& who) {
    T v = vm;
    who = kStM;
    if (vx > v) {
        v = vx;
        who = kStX;
    }
    if (vy > v) {
        v = vy;
        who = kStY;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T gap_run(uint32_t L, int32_t go, int32_t ge) {
    return L == 0 ? T(0) : T(go) + T(L - 1) * T(ge);
}

// One pair's outputs: the similarity score (int64), or in the penalty form the penalty
// -score (int32, as the reference's int32 result).
__device__ __forceinline__ void put_result(const TracebackArgs& a, uint64_t p, int64_t score,
                                           uint32_t cnt, uint32_t i, uint32_t j) {
    if (a.score32) a.score32[p] = int32_t(-score);
    else a.score[p] = score;
    a.n_ops[p] = cnt;
    a.q_begin[p] = i;
    a.r_begin[p] = j;
}

// A pair outside the host's length bounds: flagged for the host, outputs left defined.
__device__ __forceinline__ void skip_pair(const TracebackArgs& a, uint64_t p) {
    atomicOr(a.too_long, 1u);
    put_result(a, p, 0, 0, 0, 0);
}

// The residues r[0 .. n), 1 <= n <= NC, as little-endian bytes of NC / 4 words, zero past n:
// NC / 4 + 1 aligned dword loads, none behind a branch (each dword holds at least one of the
// n bytes, the ones past the last clamped to it, so no read leaves the pages the residues lie
// in), all in flight at once, then realigned with v_alignbyte.
template <uint32_t NC>
__device__ __forceinline__ void load_codes(const uint8_t* r, uint32_t n, uint32_t (&rcw)[NC / 4]) {
    const uint64_t addr = reinterpret_cast<uint64_t>(r);
    const uint64_t a0 = addr & ~uint64_t(3), last = (addr + n - 1) & ~uint64_t(3);
    const uint32_t sh = uint32_t(addr & 3u);
    uint32_t d[NC / 4 + 1];
#pragma unroll
    for (uint32_t w = 0; w <= NC / 4; ++w) {
        const uint64_t aw = a0 + 4 * w;
        d[w] = *reinterpret_cast<const uint32_t*>(aw <= last ? aw : last);
    }
#pragma unroll
    for (uint32_t w = 0; w < NC / 4; ++w) {
        const uint32_t have = n > 4 * w ? n - 4 * w : 0u;  // residues in word w
        const uint32_t keep = have >= 4 ? 0xFFFFFFFFu : (1u << (8 * have)) - 1u;
        rcw[w] = __builtin_amdgcn_alignbyte(d[w + 1], d[w], sh) & keep;
    }
}

// Dynamic LDS: the A x A similarity table, then (LDS_ROWS) the lane's reference codes
// [W][64] and the three DP rows [3][W][64]; otherwise rows in `rows_g` (3 * W * 64 T per
// workgroup) and codes read from the CSR set. dirs: (q_max + 1) * W * 64 bytes per workgroup;
// W = r_max + 1.
template <class T, bool LOCAL, bool LDS_ROWS>
__global__ __launch_bounds__(kTbLanes) void traceback_kernel(
    TracebackArgs a, T* __restrict__ rows_g, uint8_t* __restrict__ dirs_g) {
    extern __shared__ int32_t tb_lds[];
    const uint32_t lane = threadIdx.x;
    const uint32_t W = a.r_max + 1;
    int32_t* sim_l = tb_lds;
    for (uint32_t e = lane; e < a.A * a.A; e += kTbLanes) sim_l[e] = a.sim[e];
    __syncthreads();  // the only barrier: the pair loop below has none
    uint8_t* rc_l = reinterpret_cast<uint8_t*>(tb_lds + kMaxAlphabet * kMaxAlphabet);
    T* rows_l = reinterpret_cast<T*>(rc_l + ((W * kTbLanes + 15) & ~15u));
    T* rows = LDS_ROWS ? rows_l : rows_g + uint64_t(blockIdx.x) * 3 * W * kTbLanes;
    uint8_t* dirs = dirs_g + uint64_t(blockIdx.x) * (a.q_max + 1) * W * kTbLanes;
    // lane-interleaved views: row s at column j, and the direction byte of cell (i, j)
    auto R = [&](uint32_t s, uint32_t j) -> T& { return rows[(uint64_t(s) * W + j) * kTbLanes + lane]; };
    auto D = [&](uint32_t i, uint32_t j) -> uint8_t& {
        return dirs[(uint64_t(i) * W + j) * kTbLanes + lane];
    };
    T NEG;
    if constexpr (sizeof(T) == 8) NEG = T(a.neg_inf);
    else NEG = NegInf<T>::v;
    const int32_t go = a.gap_open, ge = a.gap_extend;
    const uint64_t stride = uint64_t(gridDim.x) * kTbLanes;

    for (uint64_t t = uint64_t(blockIdx.x) * kTbLanes + lane; t < a.n_pairs; t += stride) {
        const uint64_t p = a.order ? a.order[t] : t;  // shape order: lanes finish together
        const uint32_t qi = a.pair_q[p], ri = a.pair_r[p];
        const uint64_t q0 = a.q_off[qi], r0 = a.r_off[ri];
        const uint64_t m64 = a.q_off[qi + 1] - q0, n64 = a.r_off[ri + 1] - r0;
        if (m64 > a.q_max || n64 > a.r_max) {
            skip_pair(a, p);
            continue;
        }
        const uint32_t m = uint32_t(m64), n = uint32_t(n64);
        const uint8_t* q = a.q_codes + q0;
        const uint8_t* r = a.r_codes + r0;
        uint8_t* ops = a.ops + p * a.ops_stride;
        uint32_t cnt = 0, i = 0, j = 0;
        T score;
        if (m == 0 || n == 0) {
            score = LOCAL ? T(0) : -gap_run<T>(m ? m : n, go, ge);
            if (!LOCAL) {
                for (uint32_t t = 0; t < m; ++t) ops[cnt++] = 'D';
                for (uint32_t t = 0; t < n; ++t) ops[cnt++] = 'I';
            }
        } else {
            if constexpr (LDS_ROWS)  // the lane's reference codes, lane-interleaved
                for (uint32_t c = 0; c < n; ++c) rc_l[c * kTbLanes + lane] = r[c];
            // row 0
            R(0, 0) = 0;
            R(1, 0) = NEG;
            R(2, 0) = NEG;
            D(0, 0) = kStStop | (kStM << 2) | (kStM << 4);
            for (uint32_t c = 1; c <= n; ++c) {
                R(0, c) = LOCAL ? T(0) : NEG;
                R(1, c) = NEG;
                R(2, c) = LOCAL ? NEG : -gap_run<T>(c, go, ge);
                D(0, c) = kStStop | (kStM << 2) | ((c > 1 ? kStY : kStM) << 4);
            }
            T best = LOCAL ? T(0) : NEG;
            uint32_t bi = 0, bj = 0;
            for (uint32_t row = 1; row <= m; ++row) {
                // column 0 of this row; row - 1 at column 0 is the first diagonal
                T dm = R(0, 0), dx = R(1, 0), dy = R(2, 0);
                T lm = LOCAL ? T(0) : NEG;
                T lx = LOCAL ? NEG : -gap_run<T>(row, go, ge);
                T ly = NEG;
                R(0, 0) = lm;
                R(1, 0) = lx;
                R(2, 0) = ly;
                D(row, 0) = kStStop | ((row > 1 ? kStX : kStM) << 2) | (kStM << 4);
                const int32_t* srow = sim_l + uint32_t(q[row - 1]) * a.A;
                for (uint32_t c = 1; c <= n; ++c) {
                    const T um = R(0, c), ux = R(1, c), uy = R(2, c);
                    uint8_t wm, wx, wy;
                    const uint32_t rc = LDS_ROWS ? rc_l[(c - 1) * kTbLanes + lane] : r[c - 1];
                    T mv = pick(dm, dx, dy, wm) + T(srow[rc]);
                    if (LOCAL && mv < 0) {
                        mv = 0;
                        wm = kStStop;
                    }
                    const T xv = pick(T(um - go), T(ux - ge), T(uy - go), wx);
                    const T yv = pick(T(lm - go), T(lx - go), T(ly - ge), wy);
                    if (LOCAL && mv > best) {
                        best = mv;
                        bi = row;
                        bj = c;
                    }
                    R(0, c) = mv;
                    R(1, c) = xv;
                    R(2, c) = yv;
                    D(row, c) = uint8_t(wm | (wx << 2) | (wy << 4) | (mv == 0 ? kZeroBit : 0));
                    dm = um;
                    dx = ux;
                    dy = uy;
                    lm = mv;
                    lx = xv;
                    ly = yv;
                }
            }
            uint8_t st;
            bool run = true;
            if (LOCAL) {
                score = best;
                i = bi;
                j = bj;
                st = kStM;
                run = best != 0;  // no positive-scoring local alignment exists
            } else {
                score = pick(R(0, n), R(1, n), R(2, n), st);
                i = m;
                j = n;
            }
            while (run && (i > 0 || j > 0)) {
                const uint8_t dv = D(i, j);
                if (st == kStM) {
                    if (i == 0 || j == 0) break;
                    if (LOCAL && (dv & kZeroBit)) break;  // the score resets here
                    ops[cnt++] = q[i - 1] == r[j - 1] ? 'M' : 'S';
                    st = dv & 3u;
                    --i;
                    --j;
                    if (st == kStStop) break;
                } else if (st == kStX) {
                    if (i == 0) break;
                    ops[cnt++] = 'D';
                    st = (dv >> 2) & 3u;
                    --i;
                } else {
                    if (j == 0) break;
                    ops[cnt++] = 'I';
                    st = (dv >> 4) & 3u;
                    --j;
                }
            }
            // written from the end backwards: reverse into forward order
            for (uint32_t t = 0; t < cnt / 2; ++t) {
                const uint8_t x = ops[t];
                ops[t] = ops[cnt - 1 - t];
                ops[cnt - 1 - t] = x;
            }
        }
        put_result(a, p, int64_t(score), cnt, i, j);
    }
}

// Register variant for int32 cells, references of at most kTbRegCols residues and queries of
// at most kTbRegMaxQ: the three DP rows live in registers (columns 0..kTbRegCols, unrolled,
// so every index is a constant) and the lane's reference codes in eight packed registers. A
// row stops at the first 4-column group no lane of the wave reaches. The direction bytes of
// four consecutive columns go out as one dword (dwords [row][kTbRegGroups][64 lanes] per
// workgroup), and the ops are staged in LDS ([q_max + r_max][64 lanes]) and written forward
// once the walk has ended. Same recurrences, argmax order, boundary bytes and traceback rules
// as traceback_kernel. Held to three waves per SIMD (168 VGPRs, a few spilled outside the
// cell loop): 5 % faster than two waves at 187 VGPRs.
template <bool LOCAL>
__global__ __launch_bounds__(kTbLanes) __attribute__((amdgpu_waves_per_eu(3, 8))) void traceback_reg_kernel(TracebackArgs a,
                                                                 uint32_t* __restrict__ dirs_g) {
    constexpr uint32_t NC = kTbRegCols;
    constexpr uint32_t W4 = kTbRegGroups;
    extern __shared__ int32_t tb_lds[];
    const uint32_t lane = threadIdx.x;
    int32_t* sim_l = tb_lds;
    for (uint32_t e = lane; e < a.A * a.A; e += kTbLanes) sim_l[e] = a.sim[e];
    __syncthreads();  // the only barrier: the pair loop below has none
    uint8_t* ops_l = reinterpret_cast<uint8_t*>(tb_lds + kMaxAlphabet * kMaxAlphabet);
    uint32_t* dirs = dirs_g + uint64_t(blockIdx.x) * (a.q_max + 1) * W4 * kTbLanes + lane;
    const uint8_t* dbytes =
        reinterpret_cast<const uint8_t*>(dirs_g + uint64_t(blockIdx.x) * (a.q_max + 1) * W4 * kTbLanes);
    auto D = [&](uint32_t i, uint32_t j) -> uint8_t {
        return dbytes[((uint64_t(i) * W4 + (j >> 2)) * kTbLanes + lane) * 4 + (j & 3u)];
    };
    constexpr int32_t NEG = NegInf<int32_t>::v;
    const int32_t go = a.gap_open, ge = a.gap_extend;
    const uint64_t stride = uint64_t(gridDim.x) * kTbLanes;

    for (uint64_t t = uint64_t(blockIdx.x) * kTbLanes + lane; t < a.n_pairs; t += stride) {
        const uint64_t p = a.order ? a.order[t] : t;
        const uint32_t qi = a.pair_q[p], ri = a.pair_r[p];
        const uint64_t q0 = a.q_off[qi], r0 = a.r_off[ri];
        const uint64_t m64 = a.q_off[qi + 1] - q0, n64 = a.r_off[ri + 1] - r0;
        if (m64 > a.q_max || n64 > a.r_max) {  // r_max <= kTbRegCols, q_max <= kTbRegMaxQ
            skip_pair(a, p);
            continue;
        }
        const uint32_t m = uint32_t(m64), n = uint32_t(n64);
        const uint8_t* q = a.q_codes + q0;
        const uint8_t* r = a.r_codes + r0;
        uint8_t* ops = a.ops + p * a.ops_stride;
        uint32_t cnt = 0, i = 0, j = 0;
        int32_t score;
        if (m == 0 || n == 0) {
            score = LOCAL ? 0 : -gap_run<int32_t>(m ? m : n, go, ge);
            if (!LOCAL) {
                for (uint32_t t2 = 0; t2 < m; ++t2) ops[cnt++] = 'D';
                for (uint32_t t2 = 0; t2 < n; ++t2) ops[cnt++] = 'I';
            }
        } else {
            uint32_t rcw[NC / 4];
            load_codes<NC>(r, n, rcw);
            int32_t rm[NC + 1], rx[NC + 1], ry[NC + 1];  // the row above, updated in place
            rm[0] = 0;
            rx[0] = NEG;
            ry[0] = NEG;
#pragma unroll
            for (uint32_t c = 1; c <= NC; ++c) {
                rm[c] = LOCAL ? 0 : NEG;
                rx[c] = NEG;
                ry[c] = LOCAL ? NEG : -gap_run<int32_t>(c, go, ge);
            }
            // row 0's direction bytes are constants
#pragma unroll
            for (uint32_t g = 0; g < W4; ++g) {
                uint32_t v = 0;
#pragma unroll
                for (uint32_t b = 0; b < 4; ++b) {
                    const uint32_t c = 4 * g + b;
                    const uint32_t byte =
                        c == 0 ? (kStStop | (kStM << 2) | (kStM << 4))
                               : (kStStop | (kStM << 2) | ((c > 1 ? kStY : kStM) << 4));
                    v |= byte << (8 * b);
                }
                dirs[uint64_t(g) * kTbLanes] = v;
            }
            int32_t best = LOCAL ? 0 : NEG;
            uint32_t bi = 0, bj = 0;
            uint32_t qc_next = q[0];
            for (uint32_t row = 1; row <= m; ++row) {
                const int32_t* srow = sim_l + qc_next * a.A;
                if (row < m) qc_next = q[row];
                uint32_t* drow = dirs + uint64_t(row) * W4 * kTbLanes;
                int32_t dm = rm[0], dx = rx[0], dy = ry[0];
                int32_t lm = LOCAL ? 0 : NEG;
                int32_t lx = LOCAL ? NEG : -gap_run<int32_t>(row, go, ge);
                int32_t ly = NEG;
                rm[0] = lm;
                rx[0] = lx;
                ry[0] = ly;
                uint32_t acc = kStStop | ((row > 1 ? kStX : kStM) << 2) | (kStM << 4);
                uint32_t g_open = 0;  // the group acc belongs to
#pragma unroll
                for (uint32_t c = 1; c <= NC; ++c) {
                    if (c % 4 == 0) {  // columns c - 4 .. c - 1 are done
                        drow[uint64_t(c / 4 - 1) * kTbLanes] = acc;
                        acc = 0;
                        g_open = c / 4;
                        if (!__any(c <= n)) break;  // uniform: no lane has column c
                    }
                    const uint32_t rc = (rcw[(c - 1) / 4] >> (8 * ((c - 1) % 4))) & 0xFFu;
                    const int32_t um = rm[c], ux = rx[c], uy = ry[c];
                    uint8_t wm, wx, wy;
                    int32_t mv = pick(dm, dx, dy, wm) + srow[rc];
                    if (LOCAL && mv < 0) {
                        mv = 0;
                        wm = kStStop;
                    }
                    const int32_t xv = pick(um - go, ux - ge, uy - go, wx);
                    const int32_t yv = pick(lm - go, lx - go, ly - ge, wy);
                    if (LOCAL && c <= n && mv > best) {
                        best = mv;
                        bi = row;
                        bj = c;
                    }
                    rm[c] = mv;
                    rx[c] = xv;
                    ry[c] = yv;
                    acc |= uint32_t(wm | (wx << 2) | (wy << 4) | (mv == 0 ? kZeroBit : 0))
                           << (8 * (c % 4));
                    dm = um;
                    dx = ux;
                    dy = uy;
                    lm = mv;
                    lx = xv;
                    ly = yv;
                }
                drow[uint64_t(g_open) * kTbLanes] = acc;  // the last, partial group
            }
            uint8_t st;
            bool run = true;
            if (LOCAL) {
                score = best;
                i = bi;
                j = bj;
                st = kStM;
                run = best != 0;
            } else {
                int32_t fm = rm[0], fx = rx[0], fy = ry[0];
#pragma unroll
                for (uint32_t c = 1; c <= NC; ++c)
                    if (c == n) {
                        fm = rm[c];
                        fx = rx[c];
                        fy = ry[c];
                    }
                score = pick(fm, fx, fy, st);
                i = m;
                j = n;
            }
            while (run && (i > 0 || j > 0)) {
                const uint8_t dv = D(i, j);
                uint8_t op;
                if (st == kStM) {
                    if (i == 0 || j == 0) break;
                    if (LOCAL && (dv & kZeroBit)) break;
                    op = q[i - 1] == r[j - 1] ? 'M' : 'S';
                    st = dv & 3u;
                    --i;
                    --j;
                } else if (st == kStX) {
                    if (i == 0) break;
                    op = 'D';
                    st = (dv >> 2) & 3u;
                    --i;
                } else {
                    if (j == 0) break;
                    op = 'I';
                    st = (dv >> 4) & 3u;
                    --j;
                }
                ops_l[cnt * kTbLanes + lane] = op;
                ++cnt;
                if (st == kStStop) break;
            }
            for (uint32_t t2 = 0; t2 < cnt; ++t2) ops[t2] = ops_l[(cnt - 1 - t2) * kTbLanes + lane];
        }
        put_result(a, p, int64_t(score), cnt, i, j);
    }
}

// Nibble variant of the register kernel, for gap_open >= gap_extend: four bits per interior
// cell (1 <= i <= m, 1 <= j <= n) instead of a direction byte, eight cells to a dword
// ([q_max][kTbNibDwords][64 lanes] dwords per workgroup; row 0 and column 0 are not stored).
// Bits 0-1 hold the state that is the cell's own maximum in the argmax order (pick(M, X, Y)),
// or 3 for M when M = 0 in local mode (the local stop); bit 2 is set when X extends X
// (pick(M_up - go, X_up - ge, Y_up - go) picked X), bit 3 when Y extends Y. That is all the
// reference's traceback reads when gap_open >= gap_extend:
//  * M at (i, j) came from pick(M, X, Y) at (i - 1, j - 1): that cell's own bits 0-1;
//  * X at (i, j) that did not extend came from M or Y at (i - 1, j); of the two, pick takes Y
//    exactly when Y_up > M_up. The up cell's own maximum is M when M_up >= X_up, Y_up and Y
//    when Y_up > M_up, X_up, and it cannot be X: X_up > M_up, X_up >= Y_up gives
//    X_up - ge > M_up - go and X_up - ge >= Y_up - go, so X would have extended. So the
//    origin is the up cell's bits 0-1;
//  * Y likewise with the left cell (pick order M, X, then Y: X over M exactly when
//    X_left > M_left; a left cell whose maximum is Y makes Y extend).
// The walk enters state M at a cell only through one of these pointers, which then name M as
// that cell's maximum, so the local stop (M = 0) can share bits 0-1 with it; the walk's first
// cell in local mode holds the best M > 0. Row 0 and column 0 take their maxima and extension
// bits from the boundary values (global: Y along row 0 and X down column 0, each opening at
// its first step; local: M). Same recurrences, results, ops and ties as traceback_reg_kernel,
// with half its direction stores and none for row 0. The global kernel runs at two waves per
// SIMD, where its three DP rows fit in registers (at three it spills them inside the row loop,
// and is slower); the local one is faster at three.
template <bool LOCAL>
__global__ __launch_bounds__(kTbLanes) __attribute__((amdgpu_waves_per_eu(LOCAL ? 3 : 2, 8))) void traceback_nib_kernel(TracebackArgs a,
                                                                 uint32_t* __restrict__ dirs_g) {
    constexpr uint32_t NC = kTbRegCols;
    constexpr uint32_t RW = kTbNibDwords;
    extern __shared__ int32_t tb_lds[];
    const uint32_t lane = threadIdx.x;
    int32_t* sim_l = tb_lds;
    for (uint32_t e = lane; e < a.A * a.A; e += kTbLanes) sim_l[e] = a.sim[e];
    __syncthreads();  // the only barrier: the pair loop below has none
    uint8_t* ops_l = reinterpret_cast<uint8_t*>(tb_lds + kMaxAlphabet * kMaxAlphabet);
    uint32_t* dirs = dirs_g + uint64_t(blockIdx.x) * a.q_max * RW * kTbLanes + lane;
    // the four bits of interior cell (i, j), i, j >= 1
    auto nib = [&](uint32_t i, uint32_t j) -> uint32_t {
        const uint32_t w = dirs[(uint64_t(i - 1) * RW + (j - 1) / 8) * kTbLanes];
        return (w >> (4 * ((j - 1) % 8))) & 0xFu;
    };
    // the four bits of cell (i, j) when it is interior, else 0
    auto cell = [&](uint32_t i, uint32_t j) -> uint32_t { return i && j ? nib(i, j) : 0u; };
    // the state that is cell (i, j)'s maximum from its bits h, boundary cells included (3 reads
    // as M)
    auto top = [&](uint32_t i, uint32_t j, uint32_t h) -> uint8_t {
        if (i == 0 || j == 0) return LOCAL || (i == 0 && j == 0) ? kStM : i == 0 ? kStY : kStX;
        h &= 3u;
        return h == 3u ? kStM : uint8_t(h);
    };
    constexpr int32_t NEG = NegInf<int32_t>::v;
    const int32_t go = a.gap_open, ge = a.gap_extend;
    const uint64_t stride = uint64_t(gridDim.x) * kTbLanes;

    for (uint64_t t = uint64_t(blockIdx.x) * kTbLanes + lane; t < a.n_pairs; t += stride) {
        const uint64_t p = a.order ? a.order[t] : t;
        const uint32_t qi = a.pair_q[p], ri = a.pair_r[p];
        const uint64_t q0 = a.q_off[qi], r0 = a.r_off[ri];
        const uint64_t m64 = a.q_off[qi + 1] - q0, n64 = a.r_off[ri + 1] - r0;
        if (m64 > a.q_max || n64 > a.r_max) {  // r_max <= kTbRegCols, q_max <= kTbRegMaxQ
            skip_pair(a, p);
            continue;
        }
        const uint32_t m = uint32_t(m64), n = uint32_t(n64);
        const uint8_t* q = a.q_codes + q0;
        const uint8_t* r = a.r_codes + r0;
        uint8_t* ops = a.ops + p * a.ops_stride;
        uint32_t cnt = 0, i = 0, j = 0;
        int32_t score;
        if (m == 0 || n == 0) {
            score = LOCAL ? 0 : -gap_run<int32_t>(m ? m : n, go, ge);
            if (!LOCAL) {
                for (uint32_t t2 = 0; t2 < m; ++t2) ops[cnt++] = 'D';
                for (uint32_t t2 = 0; t2 < n; ++t2) ops[cnt++] = 'I';
            }
        } else {
            uint32_t rcw[NC / 4];
            load_codes<NC>(r, n, rcw);
            int32_t rm[NC + 1], rx[NC + 1], ry[NC + 1];  // the row above, updated in place
            rm[0] = 0;
            rx[0] = NEG;
            ry[0] = NEG;
#pragma unroll
            for (uint32_t c = 1; c <= NC; ++c) {
                rm[c] = LOCAL ? 0 : NEG;
                rx[c] = NEG;
                ry[c] = LOCAL ? NEG : -gap_run<int32_t>(c, go, ge);
            }
            int32_t best = LOCAL ? 0 : NEG;
            uint32_t bi = 0, bj = 0;
            uint32_t qc_next = q[0];
            for (uint32_t row = 1; row <= m; ++row) {
                const int32_t* srow = sim_l + qc_next * a.A;
                if (row < m) qc_next = q[row];
                uint32_t* drow = dirs + uint64_t(row - 1) * RW * kTbLanes;
                int32_t dm = rm[0], dx = rx[0], dy = ry[0];
                int32_t lm = LOCAL ? 0 : NEG;
                int32_t lx = LOCAL ? NEG : -gap_run<int32_t>(row, go, ge);
                int32_t ly = NEG;
                rm[0] = lm;
                rx[0] = lx;
                ry[0] = ly;
                uint32_t acc = 0, done = 0;
#pragma unroll
                for (uint32_t c = 1; c <= NC; ++c) {
                    if (c % 4 == 1 && c > 1 && !__any(c <= n)) break;  // uniform
                    const uint32_t rc = (rcw[(c - 1) / 4] >> (8 * ((c - 1) % 4))) & 0xFFu;
                    const int32_t um = rm[c], ux = rx[c], uy = ry[c];
                    int32_t mv = max(max(dm, dx), dy) + srow[rc];
                    if (LOCAL) mv = max(mv, 0);
                    const int32_t xm = um - go, xx = ux - ge, xy = uy - go;
                    const int32_t xv = max(max(xm, xx), xy);
                    const uint32_t ext_x = (xx > xm) & (xx >= xy);
                    const int32_t ym = lm - go, yx = lx - go, yy = ly - ge;
                    const int32_t yv = max(max(ym, yx), yy);
                    const uint32_t ext_y = (yy > ym) & (yy >= yx);
                    // the cell's own maximum in pick's order (M, X, Y; strictly larger wins)
                    uint32_t hd = xv > mv ? (yv > xv ? kStY : kStX) : (yv > mv ? kStY : kStM);
                    if (LOCAL && hd == kStM && mv == 0) hd = 3u;
                    if (LOCAL && c <= n && mv > best) {
                        best = mv;
                        bi = row;
                        bj = c;
                    }
                    rm[c] = mv;
                    rx[c] = xv;
                    ry[c] = yv;
                    acc |= (hd | (ext_x << 2) | (ext_y << 3)) << (4 * ((c - 1) % 8));
                    if (c % 8 == 0) {
                        drow[uint64_t(c / 8 - 1) * kTbLanes] = acc;
                        acc = 0;
                    }
                    done = c;
                    dm = um;
                    dx = ux;
                    dy = uy;
                    lm = mv;
                    lx = xv;
                    ly = yv;
                }
                if (done % 8 != 0) drow[uint64_t(done / 8) * kTbLanes] = acc;  // partial dword
            }
            uint8_t st;
            bool run = true;
            if (LOCAL) {
                score = best;
                i = bi;
                j = bj;
                st = kStM;
                run = best != 0;
            } else {
                int32_t fm = rm[0], fx = rx[0], fy = ry[0];
#pragma unroll
                for (uint32_t c = 1; c <= NC; ++c)
                    if (c == n) {
                        fm = rm[c];
                        fx = rx[c];
                        fy = ry[c];
                    }
                score = pick(fm, fx, fy, st);
                i = m;
                j = n;
            }
            // h: the bits of the walk's current cell, loaded once when the walk enters it. Each
            // step is written without branches on the state (the loads from clamped indices,
            // which m, n >= 1 keep inside the pair), so that the lanes of a wave, whose walks
            // take different turns, stay on one path.
            uint32_t h = run ? cell(i, j) : 0u;
            while (run) {
                const bool sm = st == kStM, sx = st == kStX, sy = !sm && !sx;
                if ((i == 0 && j == 0) || (sm && (i == 0 || j == 0)) || (sx && i == 0) ||
                    (sy && j == 0) || (LOCAL && sm && (h & 3u) == 3u))
                    break;  // the corner, a move off the matrix, or M = 0: the local stop
                const uint8_t qc = q[i ? i - 1 : 0], rc = r[j ? j - 1 : 0];
                const uint8_t op = sm ? (qc == rc ? 'M' : 'S') : sx ? 'D' : 'I';
                const bool ext = sx   ? (j == 0 ? i > 1 : ((h >> 2) & 1u) != 0)
                                 : sy ? (i == 0 ? j > 1 : ((h >> 3) & 1u) != 0)
                                      : false;
                i -= uint32_t(sm || sx);
                j -= uint32_t(sm || sy);
                const uint32_t w = nib(i ? i : 1u, j ? j : 1u);
                h = i && j ? w : 0u;
                st = ext ? st : top(i, j, h);
                ops_l[cnt * kTbLanes + lane] = op;
                ++cnt;
            }
            for (uint32_t t2 = 0; t2 < cnt; ++t2) ops[t2] = ops_l[(cnt - 1 - t2) * kTbLanes + lane];
        }
        put_result(a, p, int64_t(score), cnt, i, j);
    }
}

// Shape order of the pairs: a counting sort by (min(m, 31), min(n, 31)), so that the lanes of
// a wave get pairs of one shape and run the same number of DP cells and traceback steps.
// Workgroup b owns a contiguous slice of the list. Count: an LDS histogram of the slice's
// keys, written as row b of a [workgroups][bins] table. Column pass (one wave per bin): the
// counts of the slices before b, in place, and the bin's total; scan (one workgroup): where
// each bin starts. Scatter: each slice places its pairs from its own row of offsets plus the
// bin starts, with LDS atomics only. The order inside
// a bin does not matter: every output is written at the pair's own index.
constexpr uint32_t kShapeThreads = 256;

__device__ __forceinline__ uint32_t shape_key(const TracebackArgs& a, uint64_t p) {
    const uint32_t qi = a.pair_q[p], ri = a.pair_r[p];
    const uint64_t m = a.q_off[qi + 1] - a.q_off[qi], n = a.r_off[ri + 1] - a.r_off[ri];
    return uint32_t(m < 31 ? m : 31) * 32 + uint32_t(n < 31 ? n : 31);
}

__device__ __forceinline__ void shape_slice(uint64_t n, uint64_t& p0, uint64_t& p1) {
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    p0 = uint64_t(blockIdx.x) * chunk;
    p1 = p0 + chunk < n ? p0 + chunk : n;
}

__global__ __launch_bounds__(kShapeThreads) void shape_count_kernel(TracebackArgs a,
                                                                    uint32_t* __restrict__ table,
                                                                    uint32_t* __restrict__ keys) {
    __shared__ uint32_t h[kTracebackShapeBins];
    for (uint32_t b = threadIdx.x; b < kTracebackShapeBins; b += kShapeThreads) h[b] = 0;
    __syncthreads();
    uint64_t p0, p1;
    shape_slice(a.n_pairs, p0, p1);
    for (uint64_t p = p0 + threadIdx.x; p < p1; p += kShapeThreads) {
        const uint32_t k = shape_key(a, p);
        keys[p] = k;
        atomicAdd(&h[k], 1u);
    }
    __syncthreads();
    uint32_t* row = table + uint64_t(blockIdx.x) * kTracebackShapeBins;
    for (uint32_t b = threadIdx.x; b < kTracebackShapeBins; b += kShapeThreads) row[b] = h[b];
}

// Column pass: one wave per bin k turns its column of slice counts into an exclusive prefix
// over the slices (in place) and writes the bin's total.
__global__ __launch_bounds__(kShapeThreads) void shape_column_kernel(uint32_t* __restrict__ table,
                                                                     uint32_t slices,
                                                                     uint32_t* __restrict__ total) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * (kShapeThreads / 64) + (threadIdx.x >> 6);
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < slices; b0 += 64) {
        const uint32_t b = b0 + lane;
        uint32_t* cell = table + uint64_t(b) * kTracebackShapeBins + k;
        const uint32_t v = b < slices ? *cell : 0u;
        uint32_t x = v;  // inclusive wave scan
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (b < slices) *cell = carry + x - v;
        carry += __shfl(x, 63);
    }
    if (lane == 0) total[k] = carry;
}

// One workgroup of kTracebackShapeBins threads: exclusive scan of the bin totals (where each
// bin starts in the order), in place.
__global__ __launch_bounds__(kTracebackShapeBins) void shape_scan_kernel(uint32_t* __restrict__ total) {
    __shared__ uint32_t part[kTracebackShapeBins];
    const uint32_t k = threadIdx.x;
    const uint32_t own = total[k];
    part[k] = own;
    __syncthreads();
    for (uint32_t off = 1; off < kTracebackShapeBins; off <<= 1) {  // inclusive scan
        const uint32_t x = k >= off ? part[k - off] : 0u;
        __syncthreads();
        part[k] += x;
        __syncthreads();
    }
    total[k] = part[k] - own;
}

__global__ __launch_bounds__(kShapeThreads) void shape_scatter_kernel(
    uint64_t n, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ table,
    const uint32_t* __restrict__ start, uint32_t* __restrict__ order) {
    __shared__ uint32_t cur[kTracebackShapeBins];
    const uint32_t* row = table + uint64_t(blockIdx.x) * kTracebackShapeBins;
    for (uint32_t b = threadIdx.x; b < kTracebackShapeBins; b += kShapeThreads)
        cur[b] = start[b] + row[b];
    __syncthreads();
    uint64_t p0, p1;
    shape_slice(n, p0, p1);
    for (uint64_t p = p0 + threadIdx.x; p < p1; p += kShapeThreads)
        order[atomicAdd(&cur[keys[p]], 1u)] = uint32_t(p);
}

template <class T, bool LOCAL>
void launch_t(const TracebackArgs& a, bool lds_rows, uint32_t workgroups, void* rows,
              uint8_t* dirs, hipStream_t s) {
    const size_t W = size_t(a.r_max) + 1;
    const size_t sim_bytes = kMaxAlphabet * kMaxAlphabet * sizeof(int32_t);
    if constexpr (sizeof(T) == 4) {  // LDS rows for int32 only
        if (lds_rows) {
            const size_t lds = sim_bytes + ((W * kTbLanes + 15) & ~size_t(15)) +
                               3 * W * kTbLanes * sizeof(T);
            hipLaunchKernelGGL((traceback_kernel<T, LOCAL, true>), dim3(workgroups),
                               dim3(kTbLanes), lds, s, a, static_cast<T*>(rows), dirs);
            return;
        }
    }
    hipLaunchKernelGGL((traceback_kernel<T, LOCAL, false>), dim3(workgroups), dim3(kTbLanes),
                       sim_bytes, s, a, static_cast<T*>(rows), dirs);
}

}  // namespace

uint32_t traceback_order_slices(uint64_t n_pairs) {
    const uint64_t slices = (n_pairs + 4095) / 4096;
    return uint32_t(slices < kTracebackShapeSlices ? slices : kTracebackShapeSlices);
}

void launch_traceback_order(const TracebackArgs& a, uint32_t* table, uint32_t* keys,
                            uint32_t* order, hipStream_t s) {
    if (a.n_pairs == 0) return;
    const uint32_t slices = traceback_order_slices(a.n_pairs);
    hipLaunchKernelGGL(shape_count_kernel, dim3(slices), dim3(kShapeThreads), 0, s, a, table,
                       keys);
    uint32_t* total = table + uint64_t(slices) * kTracebackShapeBins;
    hipLaunchKernelGGL(shape_column_kernel, dim3(kTracebackShapeBins / (kShapeThreads / 64)),
                       dim3(kShapeThreads), 0, s, table, slices, total);
    hipLaunchKernelGGL(shape_scan_kernel, dim3(1), dim3(kTracebackShapeBins), 0, s, total);
    hipLaunchKernelGGL(shape_scatter_kernel, dim3(slices), dim3(kShapeThreads), 0, s,
                       a.n_pairs, keys, table, total, order);
}

void launch_traceback_reg(const TracebackArgs& a, uint32_t workgroups, uint32_t* dirs,
                          hipStream_t s) {
    if (a.n_pairs == 0 || workgroups == 0) return;
    const size_t lds = kMaxAlphabet * kMaxAlphabet * sizeof(int32_t) +
                       size_t(a.q_max + a.r_max) * kTbLanes;
    if (a.local)
        hipLaunchKernelGGL(traceback_reg_kernel<true>, dim3(workgroups), dim3(kTbLanes), lds, s,
                           a, dirs);
    else
        hipLaunchKernelGGL(traceback_reg_kernel<false>, dim3(workgroups), dim3(kTbLanes), lds, s,
                           a, dirs);
}

void launch_traceback_nib(const TracebackArgs& a, uint32_t workgroups, uint32_t* dirs,
                          hipStream_t s) {
    if (a.n_pairs == 0 || workgroups == 0) return;
    const size_t lds = kMaxAlphabet * kMaxAlphabet * sizeof(int32_t) +
                       size_t(a.q_max + a.r_max) * kTbLanes;
    if (a.local)
        hipLaunchKernelGGL(traceback_nib_kernel<true>, dim3(workgroups), dim3(kTbLanes), lds, s,
                           a, dirs);
    else
        hipLaunchKernelGGL(traceback_nib_kernel<false>, dim3(workgroups), dim3(kTbLanes), lds, s,
                           a, dirs);
}

void launch_traceback(const TracebackArgs& a, bool wide, bool lds_rows, uint32_t workgroups,
                      void* rows, uint8_t* dirs, hipStream_t s) {
    if (a.n_pairs == 0 || workgroups == 0) return;
    if (wide) {
        if (a.local) launch_t<int64_t, true>(a, false, workgroups, rows, dirs, s);
        else launch_t<int64_t, false>(a, false, workgroups, rows, dirs, s);
    } else {
        if (a.local) launch_t<int32_t, true>(a, lds_rows, workgroups, rows, dirs, s);
        else launch_t<int32_t, false>(a, lds_rows, workgroups, rows, dirs, s);
    }
}

}  // namespace st
