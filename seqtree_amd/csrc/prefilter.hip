// Learned prefilter for thresholded gap-block search: a small 1D-conv sequence encoder and a
// Gram-distance screen, trained and run with hand-written gfx950 kernels. There is no
// counterpart in the reference: every score the prefilter lets through is rescored by the
// exact gap-block kernel, and the exhaustive exact pass stays the judge of recall.
//
// Encoder (the sequences of a set packed one after another, each followed by two zero rows):
//   X0[r]  = emb[code(r)]                      32 channels, zero on rows outside a sequence
//   Hl[r]  = relu(sum_{k<5} X_{l-1}[r-2+k] . W_l[k] + b_l)   128 channels, zeroed likewise
//   e(s)   = mean_{p < len(s)} H3[row(s, p)]   128-d
// The model's distance is |e(a) - e(b)|^2 = s(a,a) + s(b,b) - 2 s(a,b) with s the dot
// product: one BF16 MFMA GEMM over all (query, reference) pairs with FP32 accumulation, whose
// epilogue keeps only pairs under the screening threshold.
//
// Row layout of an activation buffer: sequence s of a set whose offsets start at off[0]
// occupies rows start(s) = (off[s] - off[0]) + 2 s .. start(s) + len(s) - 1, followed by two
// zero rows that pad both its own window and the next sequence's; rows_used = start(n). The
// buffer holds R rows (rows_used rounded up to the 128-row GEMM tile, zero past rows_used)
// plus two zero guard rows before row 0 and after row R - 1, so the five-row convolution
// window of every row is in bounds. A row is "valid" when it holds a residue.
//
// Convolutions as GEMMs: the im2col row of output row r is the contiguous span
// X[(r - 2) * CIN, (r + 3) * CIN) of the row-major activations, so A is X read with row
// stride CIN and K = 5 * CIN -- no im2col buffer. The backward data pass is the same
// convolution of dY with the flipped, transposed weights.
//
// MFMA operand layout (v_mfma_f32_32x32x16_bf16): lane l supplies A[l % 32][8 (l / 32) + j]
// and B[8 (l / 32) + j][l % 32], j < 8; accumulator e of lane l is
// C[8 (e / 4) + 4 (l / 32) + e % 4][l % 32].

#include "st_internal.hpp"

#include <algorithm>

namespace st {
namespace pf {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(__bf16 v) { return float(v); }
__device__ __forceinline__ __bf16 f2bf(float v) { return __bf16(v); }

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---- embedding -----------------------------------------------------------------------
__device__ __forceinline__ uint64_t start_row(const uint64_t* off, uint32_t s) {
    return (off[s] - off[0]) + 2ull * s;
}

// One thread per (sequence, channel): the sequence's rows and its two zero rows.
__global__ __launch_bounds__(256) void embed_kernel(const uint8_t* __restrict__ codes,
                                                    const uint64_t* __restrict__ off, uint32_t n,
                                                    const __bf16* __restrict__ emb,
                                                    __bf16* __restrict__ X0,
                                                    uint8_t* __restrict__ valid) {
    const uint64_t idx = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= uint64_t(n) * kEmbDim) return;
    const uint32_t s = uint32_t(idx / kEmbDim), c = uint32_t(idx % kEmbDim);
    const uint64_t o = off[s];
    const uint32_t len = uint32_t(off[s + 1] - o);
    const uint64_t r0 = start_row(off, s);
    for (uint32_t p = 0; p < len + 2; ++p) {
        const bool ok = p < len;
        X0[(r0 + p) * kEmbDim + c] =
            ok ? emb[(uint32_t(codes[o + p]) % kEmbRows) * kEmbDim + c] : f2bf(0.f);
        if (c == 0) valid[r0 + p] = ok ? 1 : 0;
    }
}

// ---- convolution GEMM (MFMA) -----------------------------------------------------------
// Y[r][n] = sum_{kk < 5 CIN} X[(r - 2) CIN + kk] * Bt[n][kk], r < R (a multiple of 128).
// EPI 0 (forward): + bias, relu, zero on invalid rows.
// EPI 1 (backward data, hidden layer): zero where mask[r][n] <= 0 (mask = this layer's input
//        activation, itself zero on invalid rows).
// EPI 2 (backward data, embedding layer): zero on invalid rows.
constexpr uint32_t kBPad = 8;  // bf16 elements of padding per LDS row

// Software-pipelined: while the MFMAs consume K-chunk c (A fragments in registers, B staged in
// one of two LDS buffers), the global loads of chunk c + 1 are in flight; one barrier per
// chunk.
template <uint32_t CIN, uint32_t N, int EPI>
__global__ __launch_bounds__(256) void conv_kernel(const __bf16* __restrict__ X,
                                                   const __bf16* __restrict__ Bt,
                                                   const float* __restrict__ bias,
                                                   const __bf16* __restrict__ mask,
                                                   const uint8_t* __restrict__ valid,
                                                   __bf16* __restrict__ Y) {
    constexpr uint32_t K = 5 * CIN;
    constexpr uint32_t NB = N / 32;
    constexpr uint32_t KC = K % 128 == 0 ? 128 : (K % 80 == 0 ? 80 : 32);
    constexpr uint32_t NCH = K / KC;
    constexpr uint32_t KS = KC / 16;               // MFMA steps per chunk
    constexpr uint32_t PIECES = N * (KC / 8);      // 16-byte pieces of a B chunk
    constexpr uint32_t PPT = (PIECES + 255) / 256;  // pieces per thread
    static_assert(K % KC == 0 && KC % 16 == 0, "K chunking");
    __shared__ __bf16 Bs[2][N][KC + kBPad];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t li = lane & 31u, lk = lane >> 5;
    const uint64_t r0 = uint64_t(blockIdx.x) * 128 + wave * 32;
    const __bf16* arow = X + (int64_t(r0 + li) - 2) * int64_t(CIN) + 8 * lk;

    f32x16 acc[NB];
#pragma unroll
    for (uint32_t j = 0; j < NB; ++j) acc[j] = f32x16{0.f};

    bf16x8 bq[PPT], aq[KS], acur[KS];
    auto load_b = [&](uint32_t kc) {
#pragma unroll
        for (uint32_t u = 0; u < PPT; ++u) {
            const uint32_t e = tid + u * 256;
            if (e < PIECES) {
                const uint32_t n = e / (KC / 8), piece = e % (KC / 8);
                bq[u] = *reinterpret_cast<const bf16x8*>(Bt + uint64_t(n) * K + kc + piece * 8);
            }
        }
    };
    auto store_b = [&](uint32_t buf) {
#pragma unroll
        for (uint32_t u = 0; u < PPT; ++u) {
            const uint32_t e = tid + u * 256;
            if (e < PIECES) {
                const uint32_t n = e / (KC / 8), piece = e % (KC / 8);
                *reinterpret_cast<bf16x8*>(&Bs[buf][n][piece * 8]) = bq[u];
            }
        }
    };
    auto load_a = [&](uint32_t kc) {
#pragma unroll
        for (uint32_t ks = 0; ks < KS; ++ks)
            aq[ks] = *reinterpret_cast<const bf16x8*>(arow + kc + ks * 16);
    };

    load_b(0);
    load_a(0);
    store_b(0);
    __syncthreads();
    // fully unrolled, so the A fragments of consecutive chunks live in distinct registers and
    // no copy of a value still in flight forces an early wait
#pragma unroll
    for (uint32_t c = 0; c < NCH; ++c) {
#pragma unroll
        for (uint32_t ks = 0; ks < KS; ++ks) acur[ks] = aq[ks];
        if (c + 1 < NCH) {
            load_b((c + 1) * KC);
            load_a((c + 1) * KC);
        }
        const uint32_t buf = c & 1u;
#pragma unroll
        for (uint32_t ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (uint32_t j = 0; j < NB; ++j) {
                const bf16x8 b =
                    *reinterpret_cast<const bf16x8*>(&Bs[buf][j * 32 + li][ks * 16 + 8 * lk]);
                acc[j] = mfma(acur[ks], b, acc[j]);
            }
        }
        if (c + 1 < NCH) store_b(buf ^ 1u);
        __syncthreads();
    }
#pragma unroll
    for (uint32_t j = 0; j < NB; ++j) {
        const uint32_t col = j * 32 + li;
        const float bcol = EPI == 0 ? bias[col] : 0.f;
#pragma unroll
        for (uint32_t e = 0; e < 16; ++e) {
            const uint64_t row = r0 + (e / 4) * 8 + lk * 4 + (e % 4);
            float v = acc[j][e];
            if constexpr (EPI == 0) {
                v = v + bcol;
                v = (v > 0.f && valid[row]) ? v : 0.f;
            } else if constexpr (EPI == 1) {
                v = bf2f(mask[row * N + col]) > 0.f ? v : 0.f;
            } else {
                v = valid[row] ? v : 0.f;
            }
            Y[row * N + col] = f2bf(v);
        }
    }
}

// ---- mean pool ---------------------------------------------------------------------------
// One workgroup of kChannels threads per sequence. Writes the FP32 embedding, its BF16
// rounding and the squared norm of the BF16 rounding (the operand the Gram GEMM sees).
constexpr uint32_t kPoolSeqs = 4;  // sequences per pooling workgroup (kChannels threads each)

__global__ __launch_bounds__(kPoolSeqs * kChannels) void pool_kernel(
    const __bf16* __restrict__ H, const uint64_t* __restrict__ off, uint32_t n,
    float* __restrict__ ef, __bf16* __restrict__ eb, float* __restrict__ norm) {
    __shared__ float parts[kPoolSeqs][kChannels / 64];
    const uint32_t g = threadIdx.x / kChannels, c = threadIdx.x % kChannels;
    const uint32_t s = blockIdx.x * kPoolSeqs + g;
    const bool ok = s < n;  // every thread reaches the barrier below
    const uint32_t len = ok ? uint32_t(off[s + 1] - off[s]) : 0u;
    const uint64_t r0 = ok ? start_row(off, s) : 0;
    float sum = 0.f;
    for (uint32_t p = 0; p < len; ++p) sum += bf2f(H[(r0 + p) * kChannels + c]);
    const float v = len ? sum / float(len) : 0.f;
    if (ok && ef) ef[uint64_t(s) * kChannels + c] = v;
    const __bf16 vb = f2bf(v);
    if (ok && eb) eb[uint64_t(s) * kChannels + c] = vb;
    float sq = bf2f(vb) * bf2f(vb);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    if ((c & 63u) == 0) parts[g][c >> 6] = sq;
    __syncthreads();
    if (ok && c == 0 && norm) {
        float t = 0.f;
        for (uint32_t w = 0; w < kChannels / 64; ++w) t += parts[g][w];
        norm[s] = t;
    }
}

// ---- Gram screen (MFMA) --------------------------------------------------------------------
// Workgroup: 4 waves x 64 queries (two 32-row MFMA blocks per wave, operands kept in
// registers for the whole K = 128) against the references [r_begin, r_end) in tiles of 64.
// Tiles stream through a three-buffer LDS ring that all four waves share: in iteration t the
// workgroup stores tile t + 1 (loaded into registers during iteration t - 1), issues the
// global loads of tile t + 2 and runs the MFMAs of tile t, with one barrier per iteration.
// Pair (q, r) is kept when dot(q, r) >= hq[q] + hr[r], i.e. |q|^2 + |r|^2 - 2 dot <= tau
// with hq = (|q|^2 - tau) / 2 and hr = |r|^2 / 2. Kept pairs go to the wave's own LDS queue
// and are flushed to the global arrays with one global atomic per queue-full (a tile with
// more hits than the queue holds is written straight to global memory).
constexpr uint32_t kGramQ = 256;
constexpr uint32_t kTileR = 64;
constexpr uint32_t kRowPad = kChannels + 8;     // bf16 per LDS row (16-byte skew)
constexpr uint32_t kWaveQueue = 256;
constexpr uint32_t kPiecesPerThread = kTileR * kChannels / 8 / 256;  // 16-byte pieces

// Exclusive prefix sum of v over the wave and the wave total.
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, uint32_t* total) {
    const uint32_t lane = __lane_id();
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void flush_queue(const uint32_t* qq, const uint32_t* qr, uint32_t n,
                                            uint32_t* hit_q, uint32_t* hit_r, uint64_t cap,
                                            unsigned long long* count) {
    const uint32_t lane = __lane_id();
    unsigned long long g = 0;
    if (lane == 0) g = atomicAdd(count, (unsigned long long)n);
    g = __shfl(g, 0);
    wave_sync();
    for (uint32_t i = lane; i < n; i += 64)
        if (g + i < cap) {
            hit_q[g + i] = qq[i];
            hit_r[g + i] = qr[i];
        }
    wave_sync();
}

__global__ __launch_bounds__(256) void gram_kernel(const __bf16* __restrict__ Q,
                                                   const float* __restrict__ q_norm, uint32_t nq,
                                                   const __bf16* __restrict__ Rm,
                                                   const float* __restrict__ r_norm, uint32_t nr,
                                                   float tau, uint32_t r_per_group,
                                                   uint32_t* hit_q, uint32_t* hit_r, uint64_t cap,
                                                   unsigned long long* count) {
    __shared__ __bf16 ring[3][kTileR][kRowPad];
    __shared__ float ring_hr[3][kTileR];  // -|r|^2 / 2 of the tile's references (-inf past nr)
    __shared__ uint32_t queue_q[4][kWaveQueue], queue_r[4][kWaveQueue];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t li = lane & 31u, lk = lane >> 5;
    const uint32_t q0 = blockIdx.x * kGramQ + wave * 64;
    const uint32_t r_begin = blockIdx.y * r_per_group;
    const uint32_t r_end = r_begin + r_per_group < nr ? r_begin + r_per_group : nr;
    const uint32_t ntiles = r_end > r_begin ? (r_end - r_begin + kTileR - 1) / kTileR : 0;
    uint32_t* qq = queue_q[wave];
    uint32_t* qr = queue_r[wave];
    uint32_t qn = 0;  // wave-uniform queue fill
    const bool active = q0 < nq;  // waves past the last query still help stage the tiles

    bf16x8 a[2][8];
    float h[2][16];
#pragma unroll
    for (uint32_t b = 0; b < 2; ++b) {
        const uint32_t q = q0 + b * 32 + li;
#pragma unroll
        for (uint32_t ks = 0; ks < 8; ++ks)
            a[b][ks] = q < nq ? *reinterpret_cast<const bf16x8*>(Q + uint64_t(q) * kChannels +
                                                                 ks * 16 + 8 * lk)
                              : bf16x8{};
#pragma unroll
        for (uint32_t e = 0; e < 16; ++e) {
            const uint32_t qe = q0 + b * 32 + (e / 4) * 8 + lk * 4 + (e % 4);
            h[b][e] = qe < nq ? 0.5f * (q_norm[qe] - tau) : __builtin_inff();
        }
    }

    // A tile in flight: thread tid holds pieces tid + 256 u (row (tid + 256 u) / 16, piece % 16)
    // and, below kTileR, the row's raw norm (no arithmetic on a value still in flight).
    struct Stage {
        bf16x8 v[kPiecesPerThread];
        float norm;
        bool ok;
    };
    auto load_tile = [&](uint32_t t, Stage& st) {
#pragma unroll
        for (uint32_t u = 0; u < kPiecesPerThread; ++u) {
            const uint32_t e = tid + 256 * u, row = e / 16, piece = e % 16;
            const uint32_t r = r_begin + t * kTileR + row;
            st.v[u] = (t < ntiles && r < r_end)
                          ? *reinterpret_cast<const bf16x8*>(Rm + uint64_t(r) * kChannels +
                                                             piece * 8)
                          : bf16x8{};
        }
        if (tid < kTileR) {
            const uint32_t r = r_begin + t * kTileR + tid;
            st.ok = t < ntiles && r < r_end;
            st.norm = r_norm[st.ok ? r : 0];
        }
    };
    auto store_tile = [&](uint32_t buf, const Stage& st) {
#pragma unroll
        for (uint32_t u = 0; u < kPiecesPerThread; ++u) {
            const uint32_t e = tid + 256 * u, row = e / 16, piece = e % 16;
            *reinterpret_cast<bf16x8*>(&ring[buf][row][piece * 8]) = st.v[u];
        }
        if (tid < kTileR) ring_hr[buf][tid] = st.ok ? -0.5f * st.norm : -__builtin_inff();
    };
    auto compute_tile = [&](uint32_t t, uint32_t buf) {
        if (!active) return;
#pragma unroll
        for (uint32_t half = 0; half < 2; ++half) {
            // accumulators start at -hr - hq (every accumulator of a lane is in column li,
            // element e in row (e / 4) * 8 + lk * 4 + e % 4), so a hit is acc >= 0 and a
            // tile without hits is found with a max tree and one compare
            const uint32_t r = r_begin + t * kTileR + half * 32 + li;
            const float hrv = ring_hr[buf][half * 32 + li];
            f32x16 acc0, acc1;
#pragma unroll
            for (uint32_t e = 0; e < 16; ++e) {
                acc0[e] = hrv - h[0][e];
                acc1[e] = hrv - h[1][e];
            }
#pragma unroll
            for (uint32_t ks = 0; ks < 8; ++ks) {
                const bf16x8 b =
                    *reinterpret_cast<const bf16x8*>(&ring[buf][half * 32 + li][ks * 16 + 8 * lk]);
                acc0 = mfma(a[0][ks], b, acc0);
                acc1 = mfma(a[1][ks], b, acc1);
            }
            float m = acc0[0];
#pragma unroll
            for (uint32_t e = 1; e < 16; ++e) m = fmaxf(m, acc0[e]);
#pragma unroll
            for (uint32_t e = 0; e < 16; ++e) m = fmaxf(m, acc1[e]);
            if (!__any(m >= 0.f)) continue;
            uint32_t bits = 0;  // bit b * 16 + e: accumulator e of block b is a hit
#pragma unroll
            for (uint32_t e = 0; e < 16; ++e) {
                bits |= uint32_t(acc0[e] >= 0.f) << e;
                bits |= uint32_t(acc1[e] >= 0.f) << (16 + e);
            }
            uint32_t total = 0;
            const uint32_t pre = wave_scan(__popc(bits), &total);
            if (qn + total > kWaveQueue && qn) {
                flush_queue(qq, qr, qn, hit_q, hit_r, cap, count);
                qn = 0;
            }
            unsigned long long gb = 0;
            const bool direct = total > kWaveQueue;
            if (direct) {
                if (lane == 0) gb = atomicAdd(count, (unsigned long long)total);
                gb = __shfl(gb, 0);
            }
            uint32_t k = pre;
            while (bits) {
                const uint32_t bit = __ffs(bits) - 1;
                bits &= bits - 1;
                const uint32_t e = bit & 15u, blk = bit >> 4;
                const uint32_t qa = q0 + blk * 32 + (e / 4) * 8 + lk * 4 + (e % 4);
                if (!direct) {
                    qq[qn + k] = qa;
                    qr[qn + k] = r;
                } else if (gb + k < cap) {
                    hit_q[gb + k] = qa;
                    hit_r[gb + k] = r;
                }
                ++k;
            }
            if (!direct) qn += total;
        }
    };
    // Three-buffer ring: iteration t stores tile t + 1 (loaded during iteration t - 1),
    // issues the loads of tile t + 2 and computes tile t, with one barrier.
    Stage st;
    load_tile(0, st);
    store_tile(0, st);
    load_tile(1, st);
    __syncthreads();
#pragma unroll 1
    for (uint32_t t = 0; t < ntiles; ++t) {
        store_tile((t + 1) % 3, st);
        load_tile(t + 2, st);
        compute_tile(t, t % 3);
        __syncthreads();
    }
    if (qn) flush_queue(qq, qr, qn, hit_q, hit_r, cap, count);
}

// ---- training -----------------------------------------------------------------------------
// Pair loss w (d - y)^2 with d = |e_a - e_b|^2 (FP32 embeddings); gradients into dE.
__global__ __launch_bounds__(kChannels) void pair_loss_kernel(
    const float* __restrict__ ef, const uint32_t* __restrict__ pa,
    const uint32_t* __restrict__ pb, const float* __restrict__ y, const float* __restrict__ w,
    float* __restrict__ dE, float* __restrict__ loss) {
    __shared__ float part[kChannels / 64];
    const uint32_t t = blockIdx.x, c = threadIdx.x;
    const uint32_t a = pa[t], b = pb[t];
    const float diff = ef[uint64_t(a) * kChannels + c] - ef[uint64_t(b) * kChannels + c];
    float sq = diff * diff;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    if ((c & 63u) == 0) part[c >> 6] = sq;
    __syncthreads();
    float d = 0.f;
    for (uint32_t u = 0; u < kChannels / 64; ++u) d += part[u];
    const float err = d - y[t];
    const float g = 4.f * w[t] * err * diff;
    atomicAdd(&dE[uint64_t(a) * kChannels + c], g);
    atomicAdd(&dE[uint64_t(b) * kChannels + c], -g);
    if (c == 0) atomicAdd(loss, w[t] * err * err);
}

// dY3[r][c] = dE[s][c] / len where H3[r][c] > 0 on the rows of sequence s, and zero on its
// two padding rows (the rows past rows_used are zeroed separately). One thread per
// (sequence, channel).
__global__ __launch_bounds__(256) void pool_bwd_kernel(const __bf16* __restrict__ H,
                                                       const float* __restrict__ dE,
                                                       const uint64_t* __restrict__ off,
                                                       uint32_t n, __bf16* __restrict__ dY) {
    const uint64_t idx = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= uint64_t(n) * kChannels) return;
    const uint32_t s = uint32_t(idx / kChannels), c = uint32_t(idx % kChannels);
    const uint32_t len = uint32_t(off[s + 1] - off[s]);
    const uint64_t r0 = start_row(off, s);
    const float g = len ? dE[uint64_t(s) * kChannels + c] / float(len) : 0.f;
    for (uint32_t p = 0; p < len + 2; ++p) {
        const uint64_t i = (r0 + p) * kChannels + c;
        dY[i] = f2bf(p < len && bf2f(H[i]) > 0.f ? g : 0.f);
    }
}

// dW[kk][co] += sum_r X[(r - 2) CIN + kk] dY[r][co], and (blockIdx.x == 0) db[co] += sum_r
// dY[r][co], over the row slice [blockIdx.y * rows_per, ...). Tile: 64 kk x 128 co, 256
// threads x (4 kk x 8 co) accumulators.
template <uint32_t CIN>
__global__ __launch_bounds__(256) void dw_kernel(const __bf16* __restrict__ X,
                                                 const __bf16* __restrict__ dY, uint64_t R,
                                                 uint32_t rows_per, float* __restrict__ dW,
                                                 float* __restrict__ db) {
    constexpr uint32_t K = 5 * CIN;
    constexpr uint32_t TK = K % 64 == 0 ? 64 : 32;
    constexpr uint32_t RB = 16;
    __shared__ float As[RB][TK];
    __shared__ float Ds[RB][kChannels];
    const uint32_t tid = threadIdx.x;
    const uint32_t k0 = blockIdx.x * TK;
    const uint32_t tk = (tid / 16) * 4, tc = (tid % 16) * 8;  // TK / 4 = 16 groups when TK=64
    const uint64_t rs = uint64_t(blockIdx.y) * rows_per;
    const uint64_t re = rs + rows_per < R ? rs + rows_per : R;
    float acc[4][8] = {};
    float bacc = 0.f;
    for (uint64_t rb = rs; rb < re; rb += RB) {
        for (uint32_t e = tid; e < RB * TK; e += 256) {
            const uint32_t i = e / TK, kk = e % TK;
            const uint64_t r = rb + i;
            As[i][kk] = r < re ? bf2f(X[(int64_t(r) - 2) * int64_t(CIN) + k0 + kk]) : 0.f;
        }
        for (uint32_t e = tid; e < RB * kChannels; e += 256) {
            const uint32_t i = e / kChannels, c = e % kChannels;
            const uint64_t r = rb + i;
            Ds[i][c] = r < re ? bf2f(dY[r * kChannels + c]) : 0.f;
        }
        __syncthreads();
        if (tk < TK) {
#pragma unroll 4
            for (uint32_t i = 0; i < RB; ++i) {
                float av[4], dv[8];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) av[u] = As[i][tk + u];
#pragma unroll
                for (uint32_t v = 0; v < 8; ++v) dv[v] = Ds[i][tc + v];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
                    for (uint32_t v = 0; v < 8; ++v) acc[u][v] += av[u] * dv[v];
            }
        }
        if (blockIdx.x == 0 && tid < kChannels)
            for (uint32_t i = 0; i < RB; ++i) bacc += Ds[i][tid];
        __syncthreads();
    }
    if (tk < TK) {
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
            for (uint32_t v = 0; v < 8; ++v)
                atomicAdd(&dW[uint64_t(k0 + tk + u) * kChannels + tc + v], acc[u][v]);
    }
    if (blockIdx.x == 0 && tid < kChannels) atomicAdd(&db[tid], bacc);
}

// dEmb[code][c] += dX0[r][c] over valid rows; per-workgroup LDS accumulation.
__global__ __launch_bounds__(256) void embed_bwd_kernel(const __bf16* __restrict__ dX0,
                                                        const uint8_t* __restrict__ codes,
                                                        const uint64_t* __restrict__ off,
                                                        uint32_t n, uint32_t seqs_per_block,
                                                        float* __restrict__ dEmb) {
    __shared__ float acc[kMaxAlphabet * kEmbDim];
    const uint32_t tid = threadIdx.x;
    for (uint32_t e = tid; e < kMaxAlphabet * kEmbDim; e += 256) acc[e] = 0.f;
    __syncthreads();
    const uint32_t s0 = blockIdx.x * seqs_per_block;
    const uint32_t s1 = s0 + seqs_per_block < n ? s0 + seqs_per_block : n;
    for (uint32_t s = s0; s < s1; ++s) {
        const uint64_t o = off[s];
        const uint32_t len = uint32_t(off[s + 1] - o);
        const uint64_t r0 = start_row(off, s);
        for (uint32_t e = tid; e < len * kEmbDim; e += 256) {
            const uint32_t p = e / kEmbDim, c = e % kEmbDim;
            const uint64_t r = r0 + p;
            atomicAdd(&acc[(uint32_t(codes[o + p]) % kEmbRows) * kEmbDim + c],
                      bf2f(dX0[r * kEmbDim + c]));
        }
    }
    __syncthreads();
    for (uint32_t e = tid; e < kMaxAlphabet * kEmbDim; e += 256)
        if (acc[e] != 0.f) atomicAdd(&dEmb[e], acc[e]);
}

// Adam on the FP32 master parameters.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p,
                                                   const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   uint32_t n, float lr, float b1, float b2,
                                                   float c1, float c2, float eps) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 operand copies of the parameters: emb, and per layer Wt[co][kk] = W[kk][co] (forward)
// and Wf[ci][k' * 128 + co] = W[(4 - k') CIN + ci][co] (backward data).
__global__ __launch_bounds__(256) void pack_weights_kernel(const float* __restrict__ P,
                                                           __bf16* __restrict__ B) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < kEmbRows * kEmbDim) B[kOffEmbB + i] = f2bf(P[kOffEmb + i]);
    const uint32_t cin[3] = {kEmbDim, kChannels, kChannels};
    const uint32_t offW[3] = {kOffW1, kOffW2, kOffW3};
    const uint32_t offT[3] = {kOffWt1, kOffWt2, kOffWt3};
    const uint32_t offF[3] = {kOffWf1, kOffWf2, kOffWf3};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const uint32_t K = 5 * cin[l];
        if (i < K * kChannels) {
            const uint32_t kk = i / kChannels, co = i % kChannels;
            const __bf16 w = f2bf(P[offW[l] + i]);
            B[offT[l] + co * K + kk] = w;
            const uint32_t k = kk / cin[l], ci = kk % cin[l];
            B[offF[l] + uint64_t(ci) * (5 * kChannels) + (4 - k) * kChannels + co] = w;
        }
    }
}

// Entries t < n with v[t] <= cutoff, in no particular order: their index, or with pa / pb
// given the gathered (pa[t], pb[t], v[t]). Each workgroup scans a contiguous range, collects
// its entries in LDS and claims space in the output with one global atomic per LDS-full.
constexpr uint32_t kSelQueue = 1024;

__global__ __launch_bounds__(256) void select_le_kernel(const int32_t* __restrict__ v, uint64_t n,
                                                        int32_t cutoff, uint64_t per_block,
                                                        const uint32_t* __restrict__ pa,
                                                        const uint32_t* __restrict__ pb,
                                                        uint32_t* __restrict__ out_a,
                                                        uint32_t* __restrict__ out_b,
                                                        int32_t* __restrict__ out_v, uint64_t cap,
                                                        unsigned long long* __restrict__ count) {
    __shared__ uint64_t qidx[kSelQueue];
    __shared__ uint32_t qn;
    __shared__ unsigned long long base;
    const uint32_t tid = threadIdx.x;
    const uint64_t b0 = uint64_t(blockIdx.x) * per_block;
    const uint64_t b1 = b0 + per_block < n ? b0 + per_block : n;
    if (tid == 0) qn = 0;
    __syncthreads();
    auto flush = [&]() {  // all threads; qn <= kSelQueue
        const uint32_t m = qn;
        if (tid == 0) base = atomicAdd(count, (unsigned long long)m);
        __syncthreads();
        for (uint32_t i = tid; i < m; i += 256) {
            const unsigned long long k = base + i;
            const uint64_t t = qidx[i];
            if (k < cap) {
                if (pa) {
                    out_a[k] = pa[t];
                    out_b[k] = pb[t];
                    out_v[k] = v[t];
                } else {
                    out_a[k] = uint32_t(t);
                }
            }
        }
        __syncthreads();
        if (tid == 0) qn = 0;
        __syncthreads();
    };
    for (uint64_t t0 = b0; t0 < b1; t0 += 256) {
        const uint64_t t = t0 + tid;
        if (t < b1 && v[t] <= cutoff) {
            const uint32_t slot = atomicAdd(&qn, 1u);  // LDS atomic
            if (slot < kSelQueue) qidx[slot] = t;
        }
        __syncthreads();
        if (qn > kSelQueue - 256) flush();  // keep room for one more round of 256
    }
    __syncthreads();
    if (qn) flush();
}

// dst[i] = src[min(i * stride, last)] for i < count (block boundaries of an offsets array).
__global__ void gather_offsets_kernel(const uint64_t* __restrict__ src, uint32_t stride,
                                      uint32_t last, uint32_t count, uint64_t* __restrict__ dst) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = src[uint64_t(i) * stride < last ? uint64_t(i) * stride : last];
}

}  // namespace

// ---- launchers ------------------------------------------------------------------------------
void launch_gather_offsets(const uint64_t* src, uint32_t stride, uint32_t last, uint32_t count,
                           uint64_t* dst, hipStream_t s) {
    if (count == 0) return;
    hipLaunchKernelGGL(gather_offsets_kernel, dim3((count + 255) / 256), dim3(256), 0, s, src,
                       stride, last, count, dst);
}
void launch_select_le(const int32_t* v, uint64_t n, int32_t cutoff, const uint32_t* pa,
                      const uint32_t* pb, uint32_t* out_a, uint32_t* out_b, int32_t* out_v,
                      uint64_t cap, unsigned long long* count, hipStream_t s) {
    if (n == 0) return;
    // about 2048 workgroups, each a contiguous multiple of 256 entries
    const uint64_t per = std::max<uint64_t>(256, ((n + 2047) / 2048 + 255) / 256 * 256);
    const uint64_t blocks = (n + per - 1) / per;
    hipLaunchKernelGGL(select_le_kernel, dim3(uint32_t(blocks)), dim3(256), 0, s, v, n, cutoff, per,
                       pa, pb, out_a, out_b, out_v, cap, count);
}
void launch_embed(const uint8_t* codes, const uint64_t* off, uint32_t n, const __bf16* emb,
                  __bf16* X0, uint8_t* valid, hipStream_t s) {
    const uint64_t total = uint64_t(n) * kEmbDim;
    if (total == 0) return;
    hipLaunchKernelGGL(embed_kernel, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, codes,
                       off, n, emb, X0, valid);
}

void launch_conv(int layer, int epi, const __bf16* X, const __bf16* Bt, const float* bias,
                 const __bf16* mask, const uint8_t* valid, __bf16* Y, uint64_t R,
                 hipStream_t s) {
    const dim3 grid(uint32_t(R / 128)), block(256);
    if (layer == 1 && epi == 0)
        hipLaunchKernelGGL((conv_kernel<kEmbDim, kChannels, 0>), grid, block, 0, s, X, Bt, bias,
                           mask, valid, Y);
    else if (epi == 0)
        hipLaunchKernelGGL((conv_kernel<kChannels, kChannels, 0>), grid, block, 0, s, X, Bt,
                           bias, mask, valid, Y);
    else if (epi == 1)
        hipLaunchKernelGGL((conv_kernel<kChannels, kChannels, 1>), grid, block, 0, s, X, Bt,
                           bias, mask, valid, Y);
    else
        hipLaunchKernelGGL((conv_kernel<kChannels, kEmbDim, 2>), grid, block, 0, s, X, Bt, bias,
                           mask, valid, Y);
}

void launch_pool(const __bf16* H, const uint64_t* off, uint32_t n, float* ef, __bf16* eb,
                 float* norm, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(pool_kernel, dim3((n + kPoolSeqs - 1) / kPoolSeqs), dim3(kPoolSeqs * kChannels),
                       0, s, H, off, n, ef, eb, norm);
}

void launch_gram(const __bf16* Q, const float* q_norm, uint32_t nq, const __bf16* Rm,
                 const float* r_norm, uint32_t nr, float tau, uint32_t* hit_q, uint32_t* hit_r,
                 uint64_t cap, unsigned long long* count, uint32_t cus, hipStream_t s) {
    if (nq == 0 || nr == 0) return;
    const uint32_t gq = (nq + kGramQ - 1) / kGramQ;
    // reference groups: enough workgroups for a few per CU, each a whole number of tiles
    uint32_t groups = (4 * cus + gq - 1) / gq;
    const uint32_t tiles = (nr + 31) / 32;
    groups = groups < 1 ? 1 : (groups > tiles ? tiles : groups);
    const uint32_t per = ((tiles + groups - 1) / groups) * 32;
    groups = (nr + per - 1) / per;
    hipLaunchKernelGGL(gram_kernel, dim3(gq, groups), dim3(256), 0, s, Q, q_norm, nq, Rm, r_norm,
                       nr, tau, per, hit_q, hit_r, cap, count);
}

void launch_pair_loss(const float* ef, const uint32_t* pa, const uint32_t* pb, const float* y,
                      const float* w, uint32_t n_pairs, float* dE, float* loss, hipStream_t s) {
    if (n_pairs == 0) return;
    hipLaunchKernelGGL(pair_loss_kernel, dim3(n_pairs), dim3(kChannels), 0, s, ef, pa, pb, y, w, dE,
                       loss);
}

void launch_pool_bwd(const __bf16* H, const float* dE, const uint64_t* off, uint32_t n,
                     __bf16* dY, hipStream_t s) {
    const uint64_t total = uint64_t(n) * kChannels;
    if (total == 0) return;
    hipLaunchKernelGGL(pool_bwd_kernel, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, H, dE,
                       off, n, dY);
}

void launch_dw(int layer, const __bf16* X, const __bf16* dY, uint64_t R, float* dW, float* db,
               hipStream_t s) {
    const uint32_t cin = layer == 1 ? kEmbDim : kChannels;
    const uint32_t K = 5 * cin;
    const uint32_t tk = K % 64 == 0 ? 64 : 32;
    const uint32_t rows_per = 2048;
    const dim3 grid(K / tk, uint32_t((R + rows_per - 1) / rows_per));
    if (layer == 1)
        hipLaunchKernelGGL(dw_kernel<kEmbDim>, grid, dim3(256), 0, s, X, dY, R, rows_per, dW, db);
    else
        hipLaunchKernelGGL(dw_kernel<kChannels>, grid, dim3(256), 0, s, X, dY, R, rows_per, dW,
                           db);
}

void launch_embed_bwd(const __bf16* dX0, const uint8_t* codes, const uint64_t* off, uint32_t n,
                      float* dEmb, hipStream_t s) {
    if (n == 0) return;
    const uint32_t per = 64;
    hipLaunchKernelGGL(embed_bwd_kernel, dim3((n + per - 1) / per), dim3(256), 0, s, dX0, codes,
                       off, n, per, dEmb);
}

void launch_adam(float* p, const float* g, float* m, float* v, uint32_t n, float lr, float b1,
                 float b2, float c1, float c2, float eps, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, n, lr, b1,
                       b2, c1, c2, eps);
}

void launch_pack_weights(const float* P, __bf16* B, hipStream_t s) {
    const uint32_t n = 5 * kChannels * kChannels;  // the largest of the tables
    hipLaunchKernelGGL(pack_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, s, P, B);
}

}  // namespace pf
}  // namespace st
