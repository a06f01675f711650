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
    const T NEG = NegInf<T>::v;
    const int32_t go = a.gap_open, ge = a.gap_extend;
    const uint64_t stride = uint64_t(gridDim.x) * kTbLanes;

    for (uint64_t p = uint64_t(blockIdx.x) * kTbLanes + lane; p < a.n_pairs; p += stride) {
        const uint32_t qi = a.pair_q[p], ri = a.pair_r[p];
        const uint64_t q0 = a.q_off[qi], r0 = a.r_off[ri];
        const uint32_t m = uint32_t(a.q_off[qi + 1] - q0), n = uint32_t(a.r_off[ri + 1] - r0);
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
        a.score[p] = int64_t(score);
        a.n_ops[p] = cnt;
        a.q_begin[p] = i;
        a.r_begin[p] = j;
    }
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
