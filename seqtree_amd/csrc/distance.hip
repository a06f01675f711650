// Unit-cost string distances over raw bytes (reference: src/distance.cpp:68-95, driven by
// src/distance.cpp:29-64).
//
// Hamming: every string has the same length (checked before launch), so each thread keeps
// one reference's words in registers and walks a block of queries whose words are
// wave-uniform (scalar loads). Per 4 bytes: x = a ^ b, then the high bit of each byte of
// ((x & 0x7f7f7f7f) + 0x7f7f7f7f) | x marks a non-zero byte; one popcount counts them.
// The kernel is bound by writing the int32 matrix to HBM.
//
// Levenshtein: Myers' bit-parallel algorithm (Myers 1999, in Hyyro's formulation) with the
// query as the pattern -- its match masks Peq[byte] sit in LDS -- and one reference per lane
// as the text. Global edit distance injects a +1 horizontal delta at row 0 on every column.
// Patterns longer than 64 use the block-carry form (Hyyro 2003): each 64-row block passes
// its bottom horizontal delta (+1/0/-1) to the next. Queries longer than 256 fall back to
// the plain two-row DP, one pair per thread.

#include "st_internal.hpp"

namespace st {
namespace {

__global__ void pack_rows_kernel(const uint8_t* __restrict__ bytes,
                                 const uint64_t* __restrict__ off, uint64_t n, uint32_t nwords,
                                 uint32_t* __restrict__ words) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * nwords) return;
    const uint64_t i = e / nwords;
    const uint32_t w = uint32_t(e - i * nwords);
    const uint64_t o = off[i];
    const uint64_t L = off[i + 1] - o;
    uint32_t word = 0;
    for (uint32_t b = 0; b < 4; ++b) {
        const uint64_t p = 4ull * w + b;
        if (p < L) word |= uint32_t(bytes[o + p]) << (8 * b);
    }
    words[e] = word;
}

constexpr uint32_t kHamQueries = 32;  // queries per thread (rows of the output per workgroup)

template <int NW>
__global__ __launch_bounds__(256) void hamming_kernel(DistanceArgs a,
                                                      const uint32_t* __restrict__ aw,
                                                      uint32_t nwords, uint32_t nbx) {
    const uint32_t bx = blockIdx.x % nbx;
    const uint64_t i0 = uint64_t(blockIdx.x / nbx) * kHamQueries;
    const uint64_t k = uint64_t(bx) * 256 + threadIdx.x;
    if (k >= a.n_b) return;
    const uint32_t npad = a.refs.npad;
    const uint32_t* rw_ptr = a.refs.words + k;
    uint32_t rw[NW > 0 ? NW : 1];
    if constexpr (NW > 0) {
#pragma unroll
        for (int w = 0; w < NW; ++w) rw[w] = uint32_t(w) < nwords ? rw_ptr[uint64_t(w) * npad] : 0u;
    }
    const uint64_t iend = i0 + kHamQueries < a.n_a ? i0 + kHamQueries : a.n_a;
    for (uint64_t i = i0; i < iend; ++i) {
        const uint32_t* qa = aw + i * nwords;
        uint32_t acc = 0;
        if constexpr (NW > 0) {
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                if (uint32_t(w) < nwords) {
                    const uint32_t x = qa[w] ^ rw[w];
                    const uint32_t t = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
                    acc += __popc(t & 0x80808080u);
                }
            }
        } else {
            for (uint32_t w = 0; w < nwords; ++w) {
                const uint32_t x = qa[w] ^ rw_ptr[uint64_t(w) * npad];
                const uint32_t t = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
                acc += __popc(t & 0x80808080u);
            }
        }
        a.out[i * a.n_b + k] = int32_t(acc);
    }
}

// Same computation, four consecutive references per thread: each query row is written with
// one 16-byte nontemporal store per thread (1 KiB per wave instruction), which the HBM write
// stream prefers to 4-byte stores. Requires n_b % 4 == 0 so every row start is 16-byte aligned.
template <int NW>
__global__ __launch_bounds__(256) void hamming4_kernel(DistanceArgs a,
                                                       const uint32_t* __restrict__ aw,
                                                       uint32_t nwords, uint32_t nbx) {
    const uint32_t bx = blockIdx.x % nbx;
    const uint64_t i0 = uint64_t(blockIdx.x / nbx) * kHamQueries;
    const uint64_t k0 = (uint64_t(bx) * 256 + threadIdx.x) * 4;
    if (k0 >= a.n_b) return;
    const uint32_t npad = a.refs.npad;
    uint32_t rw[4][NW];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int w = 0; w < NW; ++w)
            rw[c][w] = uint32_t(w) < nwords ? a.refs.words[uint64_t(w) * npad + k0 + c] : 0u;
    const uint64_t iend = i0 + kHamQueries < a.n_a ? i0 + kHamQueries : a.n_a;
    for (uint64_t i = i0; i < iend; ++i) {
        const uint32_t* qa = aw + i * nwords;
        uint32_t acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (uint32_t(w) < nwords) {
                const uint32_t qw = qa[w];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint32_t x = qw ^ rw[c][w];
                    const uint32_t t = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
                    acc[c] += __popc(t & 0x80808080u);
                }
            }
        }
        typedef int v4i __attribute__((ext_vector_type(4)));
        const v4i v = {int(acc[0]), int(acc[1]), int(acc[2]), int(acc[3])};
        __builtin_nontemporal_store(v, reinterpret_cast<v4i*>(a.out + i * a.n_b + k0));
    }
}

__global__ void hamming_check_kernel(const uint64_t* __restrict__ a_off, uint64_t n_a,
                                     const uint64_t* __restrict__ b_off, uint64_t n_b,
                                     unsigned long long* __restrict__ res) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t la0 = a_off[1] - a_off[0];
    const uint64_t lb0 = b_off[1] - b_off[0];
    if (t < n_b && b_off[t + 1] - b_off[t] != la0) atomicMin(&res[0], (unsigned long long)t);
    if (t < n_a && a_off[t + 1] - a_off[t] != lb0) atomicMin(&res[1], (unsigned long long)t);
}

template <int W>
__global__ __launch_bounds__(kThreads) void lev_myers_kernel(DistanceArgs a,
                                                             const uint32_t* __restrict__ list,
                                                             uint64_t base) {
    __shared__ unsigned long long peq[256 * W];
    __shared__ int32_t stage[kChunk];

    const uint32_t tid = threadIdx.x;
    const uint32_t nch = a.refs.nchunks;
    const uint32_t c = blockIdx.x % nch;
    const uint32_t qi = list[base + blockIdx.x / nch];
    const uint64_t q0 = a.a_off[qi];
    const uint32_t m = uint32_t(a.a_off[qi + 1] - q0);
    const uint8_t* q = a.a_bytes + q0;

    for (uint32_t e = tid; e < 256u * W; e += kThreads) peq[e] = 0ull;
    __syncthreads();
    for (uint32_t p = tid; p < m; p += kThreads)
        atomicOr(&peq[uint32_t(q[p]) * W + (p >> 6)], 1ull << (p & 63u));
    __syncthreads();

    const uint32_t lw = m ? (m - 1) >> 6 : 0;  // last pattern block
    const uint32_t hib = m ? (m - 1) & 63u : 0;
    const uint32_t npad = a.refs.npad;

    for (uint32_t it = 0; it < kSlotsPerThread; ++it) {
        const uint32_t slot = c * kChunk + it * kThreads + tid;
        const uint32_t k = a.refs.orig[slot];
        if (k == kNoSlot) continue;
        const uint32_t n = a.refs.len[slot];
        int32_t res;
        if (m == 0) {
            res = int32_t(n);
        } else {
            unsigned long long Pv[W], Mv[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                Pv[w] = ~0ull;
                Mv[w] = 0ull;
            }
            int32_t score = int32_t(m);
            const uint32_t* wp = a.refs.words + slot;
            uint32_t word = 0;
            for (uint32_t j = 0; j < n; ++j) {
                if ((j & 3u) == 0) word = wp[uint64_t(j >> 2) * npad];
                const uint32_t ch = (word >> (8 * (j & 3u))) & 0xFFu;
                const unsigned long long* pe = peq + ch * W;
                unsigned long long hin_p = 1ull, hin_n = 0ull;  // +1 from row 0
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    if (uint32_t(w) <= lw) {
                        unsigned long long eq = pe[w];
                        const unsigned long long pv = Pv[w], mvv = Mv[w];
                        const unsigned long long xv = eq | mvv;
                        eq |= hin_n;
                        const unsigned long long xh = (((eq & pv) + pv) ^ pv) | eq;
                        unsigned long long ph = mvv | ~(xh | pv);
                        unsigned long long mh = pv & xh;
                        if (uint32_t(w) == lw) {
                            score += int32_t((ph >> hib) & 1ull);
                            score -= int32_t((mh >> hib) & 1ull);
                        }
                        const unsigned long long hout_p = ph >> 63, hout_n = mh >> 63;
                        ph = (ph << 1) | hin_p;
                        mh = (mh << 1) | hin_n;
                        Pv[w] = mh | ~(xv | ph);
                        Mv[w] = ph & xv;
                        hin_p = hout_p;
                        hin_n = hout_n;
                    }
                }
            }
            res = score;
        }
        stage[k - c * kChunk] = res;
    }
    __syncthreads();

    const uint64_t col0 = uint64_t(c) * kChunk;
    const uint64_t lim64 = a.n_b - col0;
    const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
    int32_t* row = a.out + uint64_t(qi) * a.n_b + col0;
    for (uint32_t t = tid; t < lim; t += kThreads) row[t] = stage[t];
}

// Two-row unit-cost DP with the row over the shorter string (reference distance.cpp:78-95).
__device__ int32_t lev_dp(const uint8_t* x, uint32_t m, const uint8_t* y, uint32_t n,
                          int32_t* buf) {
    if (n > m) {
        const uint8_t* t = x;
        x = y;
        y = t;
        const uint32_t tm = m;
        m = n;
        n = tm;
    }
    if (n == 0) return int32_t(m);
    int32_t* prev = buf;
    int32_t* cur = buf + (n + 1);
    for (uint32_t j = 0; j <= n; ++j) prev[j] = int32_t(j);
    for (uint32_t i = 1; i <= m; ++i) {
        cur[0] = int32_t(i);
        const uint8_t xi = x[i - 1];
        for (uint32_t j = 1; j <= n; ++j) {
            const int32_t cost = xi == y[j - 1] ? 0 : 1;
            cur[j] = min(min(prev[j] + 1, cur[j - 1] + 1), prev[j - 1] + cost);
        }
        int32_t* t = prev;
        prev = cur;
        cur = t;
    }
    return prev[n];
}

__global__ __launch_bounds__(256) void lev_generic_kernel(DistanceArgs a,
                                                          const uint32_t* __restrict__ list,
                                                          uint32_t count,
                                                          int32_t* __restrict__ scratch,
                                                          uint64_t stride) {
    const uint64_t slots = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int32_t* buf = scratch + t * 2 * stride;
    const uint64_t total = uint64_t(count) * a.n_b;
    for (uint64_t p = t; p < total; p += slots) {
        const uint32_t qi = list[p / a.n_b];
        const uint64_t k = p % a.n_b;
        const uint64_t x0 = a.a_off[qi], y0 = a.b_off[k];
        a.out[uint64_t(qi) * a.n_b + k] =
            lev_dp(a.a_bytes + x0, uint32_t(a.a_off[qi + 1] - x0), a.b_bytes + y0,
                   uint32_t(a.b_off[k + 1] - y0), buf);
    }
}

template <int W>
void launch_myers(const DistanceArgs& a, const uint32_t* list, uint32_t count, hipStream_t s) {
    const uint64_t nch = a.refs.nchunks;
    const uint64_t max_q = (uint64_t(1) << 30) / nch;
    for (uint64_t q0 = 0; q0 < count; q0 += max_q) {
        const uint64_t nq = count - q0 < max_q ? count - q0 : max_q;
        hipLaunchKernelGGL((lev_myers_kernel<W>), dim3(uint32_t(nq * nch)), dim3(kThreads), 0, s,
                           a, list, q0);
    }
}

}  // namespace

void launch_pack_rows(const uint8_t* bytes, const uint64_t* off, uint64_t n, uint32_t nwords,
                      uint32_t* words, hipStream_t s) {
    const uint64_t total = n * nwords;
    if (total == 0) return;
    hipLaunchKernelGGL(pack_rows_kernel, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s,
                       bytes, off, n, nwords, words);
}

void launch_hamming(const DistanceArgs& a, const uint32_t* a_words, uint32_t nwords,
                    hipStream_t s) {
    if (a.n_a == 0 || a.n_b == 0) return;
    const bool vec4 = a.n_b % 4 == 0 && nwords <= 4;
    const uint32_t nbx = uint32_t((a.n_b + (vec4 ? 1023 : 255)) / (vec4 ? 1024 : 256));
    const uint64_t nby = (a.n_a + kHamQueries - 1) / kHamQueries;
    // one launch covers up to 2^30 workgroups; split the query blocks beyond that
    const uint64_t max_by = (uint64_t(1) << 30) / nbx;
    for (uint64_t by0 = 0; by0 < nby; by0 += max_by) {
        const uint64_t cnt = nby - by0 < max_by ? nby - by0 : max_by;
        DistanceArgs sub = a;
        sub.out = a.out + by0 * kHamQueries * a.n_b;
        sub.n_a = a.n_a - by0 * kHamQueries;
        const uint32_t* aw = a_words + by0 * kHamQueries * nwords;
        const dim3 grid(uint32_t(cnt * nbx));
        if (vec4)
            hipLaunchKernelGGL(hamming4_kernel<4>, grid, dim3(256), 0, s, sub, aw, nwords, nbx);
        else if (nwords <= 4)
            hipLaunchKernelGGL(hamming_kernel<4>, grid, dim3(256), 0, s, sub, aw, nwords, nbx);
        else if (nwords <= 16)
            hipLaunchKernelGGL(hamming_kernel<16>, grid, dim3(256), 0, s, sub, aw, nwords, nbx);
        else
            hipLaunchKernelGGL(hamming_kernel<0>, grid, dim3(256), 0, s, sub, aw, nwords, nbx);
    }
}

void launch_hamming_check(const uint64_t* a_off, uint64_t n_a, const uint64_t* b_off,
                          uint64_t n_b, uint64_t* res, hipStream_t s) {
    const uint64_t n = n_a > n_b ? n_a : n_b;
    if (n == 0) return;
    hipLaunchKernelGGL(hamming_check_kernel, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s,
                       a_off, n_a, b_off, n_b, reinterpret_cast<unsigned long long*>(res));
}

void launch_levenshtein_myers(const DistanceArgs& a, int words, const uint32_t* list,
                              uint32_t count, hipStream_t s) {
    if (count == 0 || a.refs.nchunks == 0) return;
    switch (words) {
        case 1: launch_myers<1>(a, list, count, s); break;
        case 2: launch_myers<2>(a, list, count, s); break;
        case 3: launch_myers<3>(a, list, count, s); break;
        default: launch_myers<4>(a, list, count, s); break;
    }
}

void launch_levenshtein_generic(const DistanceArgs& a, const uint32_t* list, uint32_t count,
                                int32_t* scratch, uint32_t slots, uint64_t stride,
                                hipStream_t s) {
    if (count == 0 || a.n_b == 0) return;
    hipLaunchKernelGGL(lev_generic_kernel, dim3(slots / 256), dim3(256), 0, s, a, list, count,
                       scratch, stride);
}

}  // namespace st
