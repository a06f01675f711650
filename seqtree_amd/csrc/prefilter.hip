// Learned prefilter for thresholded gap-block search: a small 1D-conv sequence encoder and a
// Gram-distance screen, trained and run with hand-written gfx950 kernels. There is no
// counterpart in the reference: every score the prefilter lets through is rescored by the
// exact gap-block kernel, and the exhaustive exact pass stays the measure of recall.
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
#include <cstddef>

namespace st {
namespace pf {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(__bf16 v) { return float(v); }
__device__ __forceinline__ __bf16 f2bf(float v) { return __bf16(v); }

// Buffer resource word 3 for raw (untyped) loads on gfx9-family parts.
constexpr int kBufferFlags = 0x00020000;

typedef __attribute__((address_space(3))) void* lds_void;

// s_waitcnt vmcnt(N), leaving the other counters alone (gfx9 field layout).
template <uint32_t N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N < 64, "vmcnt holds six bits");
    __builtin_amdgcn_s_waitcnt((N & 0xFu) | ((N >> 4) << 14) | (0x7u << 4) | (0xFu << 8));
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---- embedding -----------------------------------------------------------------------
constexpr uint32_t kEncTile = 128;  // rows per workgroup of the fused encoder
__device__ __forceinline__ uint64_t start_row(const uint64_t* off, uint32_t s) {
    return (off[s] - off[0]) + 2ull * s;
}

// Four threads per sequence, one per 16-byte piece (8 channels) of the 64-byte rows: the
// sequence's rows and its two zero rows, with the table in LDS. With tile_first, also records
// for every 128-row tile whose first row falls in this sequence's span (its rows and two zero
// rows; spans partition the used rows) the sequence index.
constexpr uint32_t kEmbPieces = kEmbDim / 8;
__global__ __launch_bounds__(256) void embed_kernel(const uint8_t* __restrict__ codes,
                                                    const uint64_t* __restrict__ off, uint32_t n,
                                                    const __bf16* __restrict__ emb,
                                                    __bf16* __restrict__ X0,
                                                    uint8_t* __restrict__ valid,
                                                    uint32_t* __restrict__ tile_first) {
    __shared__ bf16x8 table[kEmbRows * kEmbPieces];
    for (uint32_t e = threadIdx.x; e < kEmbRows * kEmbPieces; e += 256)
        table[e] = reinterpret_cast<const bf16x8*>(emb)[e];
    __syncthreads();
    const uint64_t idx = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= uint64_t(n) * kEmbPieces) return;
    const uint32_t s = uint32_t(idx / kEmbPieces), piece = uint32_t(idx % kEmbPieces);
    const uint64_t o = off[s];
    const uint32_t len = uint32_t(off[s + 1] - o);
    const uint64_t r0 = start_row(off, s);
    if (tile_first && piece == 0) {
        const uint64_t t = (r0 + kEncTile - 1) / kEncTile;  // spans are shorter than a tile
        if (t * kEncTile < r0 + len + 2) tile_first[t] = s;
    }
    bf16x8* dst = reinterpret_cast<bf16x8*>(X0) + r0 * kEmbPieces + piece;
    const bf16x8* tp = table + piece;
#pragma unroll 4
    for (uint32_t p = 0; p < len; ++p)
        dst[uint64_t(p) * kEmbPieces] = tp[(uint32_t(codes[o + p]) % kEmbRows) * kEmbPieces];
    dst[uint64_t(len) * kEmbPieces] = bf16x8{};
    dst[uint64_t(len + 1) * kEmbPieces] = bf16x8{};
    // the valid bytes: each of the four threads takes every fourth row
    for (uint32_t p = piece; p < len + 2; p += kEmbPieces) valid[r0 + p] = p < len ? 1 : 0;
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

// ---- fused encoder (inference) -------------------------------------------------------------
// One workgroup computes H3 for the 128 rows [r0, r0 + 128) straight from X0 and mean-pools it,
// without writing H1, H2 or H3 to global memory: conv1 over the 136 rows [r0 - 4, r0 + 132),
// conv2 over the 132 rows [r0 - 2, r0 + 130), conv3 over the 128 rows, each layer's output
// rounded to BF16 as the unfused kernels store it. Wave w computes output channels
// [32 w, 32 w + 32) of every 32-row block (five for conv1 and conv2, whose rows past the needed
// ones are clamped and discarded, four for conv3); its weight fragments come from global memory
// (L2) into a register ring eight k-steps ahead, the activation fragments from LDS one k-step
// ahead. conv1 and conv2 take the weights as the MFMA's first operand, so each lane's
// accumulators hold four consecutive channels of one row per group and the output goes to LDS
// as 8-byte stores; conv3 takes the activations first, and its accumulators feed the pooling
// MFMAs (enc_pool) without leaving registers.
//
// Frames: output frame row j of every layer reads its input frame rows j .. j + 4 (the input
// frame starts two rows earlier). Frames start at global rows r0 - 6 (X0), r0 - 4 (H1),
// r0 - 2 (H2) and r0 (H3).
constexpr uint32_t kEncRows = kEncTile;
constexpr uint32_t kXStride = 80;   // bytes per X0 row in LDS: 64 + 16 (conflict-free reads)
constexpr uint32_t kHStride = 272;  // bytes per H row in LDS: 256 + 16
constexpr uint32_t kXRows = kEncRows + 12, kH1Rows = kEncRows + 8, kH2Rows = kEncRows + 4;
// H1 (136 rows), later the staging of the pooled embeddings; X0, later H2 (132 rows plus 4 rows
// that the last block's discarded rows land in)
constexpr uint32_t kEncRegionA = kH1Rows * kHStride;
constexpr uint32_t kEncRegionB = (kH2Rows + 4) * kHStride;
constexpr uint32_t kEncMaskRows = 160;  // per-layer valid masks, one byte per frame row
static_assert(kXRows * kXStride <= kEncRegionB && kEncRows * kHStride <= kEncRegionA, "LDS");

// The lane and wave indices of the fused encoder's per-tile code, recomputed at each use: the
// empty asm makes the lane index a new value each time, so the addresses derived from it are not
// hoisted out of the tile loop (where they would stay live across every layer and spill), and
// the wave index is a scalar.
__device__ __forceinline__ uint32_t enc_lane() {
    uint32_t l = __lane_id();
    asm volatile("" : "+v"(l));
    return l;
}
__device__ __forceinline__ uint32_t enc_wave() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Weight fragments in flight per lane. The fused encoder streams the weight fragments of conv1
// (10 k-steps), conv2 (40) and conv3 (40), then the next tile's conv1, through one register
// ring: the fragment of stream position f sits in slot f % kRingDepth, and every layer's
// k-step count is a multiple of the depth, so slot ks % kRingDepth holds k-step ks of the
// current layer in each one.
constexpr uint32_t kRingDepth = 5;

// Block b of output channels [32 w, 32 w + 32) and rows 32 b + (0 .. 31) (rows past `clamp` read
// row `clamp`): the sum over K = 5 CIN of A[row][k] W[k][channel], A row j of the input frame
// span j .. j + 4 in LDS (`in`, row stride `stride` bytes). ROWS puts the rows in the
// accumulators' rows (A as the MFMA's first operand), otherwise the channels. On entry the ring
// holds this layer's first kRingDepth weight fragments (Wt); k-step ks consumes its slot and
// refills it with the fragment kRingDepth k-steps further on in the stream, which past this
// layer's last k-step is the next layer's (Wn, CIN_NEXT input channels). The accumulators start
// from `init` (the layer's bias in the accumulator layout), taken as the first MFMA's addend.
template <uint32_t CIN, uint32_t NBLK, bool ROWS, uint32_t CIN_NEXT>
__device__ __forceinline__ void enc_layer(const char* in, uint32_t stride, uint32_t clamp,
                                          const __bf16* __restrict__ Wt,
                                          const __bf16* __restrict__ Wn, const f32x16& init,
                                          bf16x8 (&ring)[kRingDepth], f32x16 (&acc)[NBLK]) {
    constexpr uint32_t K = 5 * CIN, KS = K / 16, P = kRingDepth, KN = 5 * CIN_NEXT;
    static_assert(KS % P == 0 && KN / 16 >= P, "ring slots follow every layer's k-steps");
    const uint32_t lane = enc_lane(), li = lane & 31u, lk = lane >> 5, w = enc_wave();
    const __bf16* wrow = Wt + uint64_t(32 * w + li) * K + 8 * lk;
    const __bf16* wnext = Wn + uint64_t(32 * w + li) * KN + 8 * lk;
    const char* arow[NBLK];
#pragma unroll
    for (uint32_t b = 0; b < NBLK; ++b) {
        const uint32_t j = 32 * b + li;
        arow[b] = in + (j < clamp ? j : clamp) * stride + 16 * lk;
    }
    // the activation fragments of k-step ks + 1 are read while the MFMAs of k-step ks run: the
    // scheduling barriers keep the compiler from sinking each read to just before its MFMA,
    // which would expose the LDS latency once per MFMA
    const auto a_frag = [&](uint32_t ks, uint32_t b) {
        const uint32_t kk = (16 * ks) / CIN, ci = (16 * ks) % CIN;
        return *reinterpret_cast<const bf16x8*>(arow[b] + kk * stride + ci * 2);
    };
    bf16x8 cur[NBLK];
#pragma unroll
    for (uint32_t b = 0; b < NBLK; ++b) cur[b] = a_frag(0, b);
#pragma unroll
    for (uint32_t ks = 0; ks < KS; ++ks) {
        const bf16x8 next = *reinterpret_cast<const bf16x8*>(
            ks + P < KS ? wrow + 16 * (ks + P) : wnext + 16 * (ks + P - KS));
        asm volatile("" ::: "memory");  // issue the weight load here
        bf16x8 nxt[NBLK];
#pragma unroll
        for (uint32_t b = 0; b < NBLK; ++b) nxt[b] = ks + 1 < KS ? a_frag(ks + 1, b) : cur[b];
        __builtin_amdgcn_sched_barrier(0);
        const bf16x8 bw = ring[ks % P];
#pragma unroll
        for (uint32_t b = 0; b < NBLK; ++b) {
            const f32x16& c = ks == 0 ? init : acc[b];
            acc[b] = ROWS ? mfma(cur[b], bw, c) : mfma(bw, cur[b], c);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (uint32_t b = 0; b < NBLK; ++b) cur[b] = nxt[b];
        ring[ks % P] = next;
    }
}

// The bias of a layer whose accumulators hold channels in their rows (lane (li, lk), element
// 4 g + u: channel 32 w + 8 g + 4 lk + u), or one channel per lane (32 w + li).
__device__ __forceinline__ f32x16 enc_bias_channels(const float* bias) {
    const uint32_t lane = enc_lane(), lk = lane >> 5, w = enc_wave();
    f32x16 v;
#pragma unroll
    for (uint32_t g = 0; g < 4; ++g) {
        const float4 b4 = *reinterpret_cast<const float4*>(bias + 32 * w + 8 * g + 4 * lk);
        v[4 * g] = b4.x;
        v[4 * g + 1] = b4.y;
        v[4 * g + 2] = b4.z;
        v[4 * g + 3] = b4.w;
    }
    return v;
}
__device__ __forceinline__ f32x16 enc_bias_lane(const float* bias) {
    const float b = bias[32 * enc_wave() + (enc_lane() & 31u)];
    f32x16 v;
#pragma unroll
    for (uint32_t e = 0; e < 16; ++e) v[e] = b;
    return v;
}

// ReLU and the frame's valid mask in one step: `cap` is +inf on a valid row and 0 on any
// other, so med3(v, 0, cap) is max(v, 0) on valid rows and 0 elsewhere.
__device__ __forceinline__ float enc_relu(float v, float cap) {
    return __builtin_amdgcn_fmed3f(v, 0.f, cap);
}

// The accumulators (bias included), after ReLU and the frame's valid mask (`cap`, one float
// per frame row), rounded to BF16 into `out` (LDS, row stride kHStride); the last block
// stores its first 8 LAST rows only (the rows past those are not needed). Lane (li, lk) holds
// row 32 b + li of block b and, in group g, the channels 32 w + 8 g + 4 lk + u, u < 4: one
// 8-byte store per group.
template <uint32_t NBLK, uint32_t LAST>
__device__ __forceinline__ void enc_store(const f32x16 (&acc)[NBLK], const float* cap,
                                          char* out) {
    const uint32_t lane = enc_lane(), li = lane & 31u, lk = lane >> 5, w = enc_wave();
    const uint32_t c0 = 32 * w + 4 * lk;
#pragma unroll
    for (uint32_t b = 0; b < NBLK; ++b) {
        const uint32_t j = 32 * b + li;
        if (b + 1 == NBLK && li >= 8 * LAST) continue;
        const float hi = cap[j];
        char* dst = out + j * kHStride + c0 * 2;
#pragma unroll
        for (uint32_t g = 0; g < 4; ++g) {
            bf16x4 pk;
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) pk[u] = f2bf(enc_relu(acc[b][4 * g + u], hi));
            *reinterpret_cast<bf16x4*>(dst + 16 * g) = pk;
        }
    }
}

// Pooling records of the fused encoder, per tile t: the FP32 channel sums of the sequence that
// starts in tile t and runs past its end (`head`, with its index in head_seq, or kNoSeq) and
// of the rows in tile t of the sequence that started before it (`tail`).
constexpr uint32_t kNoSeq = 0xFFFFFFFFu;
constexpr uint32_t kEncMaxSeqs = kEncRows / 2 + 1;  // sequences meeting one tile, at most

struct EncPool {
    const uint64_t* off;       // the block's offsets, off[0 .. n]
    uint32_t n;
    const uint32_t* tile_first;  // per tile: the sequence whose span holds its first row, or n
    __bf16* eb;                // per sequence: the BF16 embedding
    float* norm;               // per sequence: squared norm of eb
    float* part;               // per tile: head sums [128], tail sums [128]
    uint32_t* head_seq;        // per tile
};

// The sequences s0 + j, j < cnt, that meet a tile: each one's first residue row relative to the
// tile's first row (negative for the one that started before it), its length and the
// reciprocal of its length (0 for an empty one). Entries past cnt are stale.
struct EncSeqs {
    int32_t lo[96];
    uint32_t len[96];
    float inv[96];
};
static_assert(96 >= (kEncMaxSeqs + 31) / 32 * 32, "three blocks of 32 sequences");
constexpr uint32_t kStageStride = 136;  // floats per staged row of sums (conflict-free stores)
static_assert(32 * kStageStride * 4 <= kEncRegionB, "the sums' stage fits the H2 region");

// Mean pool of the tile from conv3's accumulators, as MFMAs: sums[s][c] = sum over the tile's
// rows r of M[s][r] H3[r][c], M the 0/1 membership of row r in sequence s's residue rows.
// conv3 runs with the activations as the first MFMA operand, so lane (li, lk) of wave w holds
// channel c = 32 w + li of the rows 32 b + 8 g + 4 lk + u (element 4 g + u of block b), bias
// included. After ReLU, the frame mask (`cap`) and the BF16 rounding (the values the unfused
// layers store),
// elements 8 h .. 8 h + 7 of block b are the lane's second operand for k-step 2 b + h, k slot i
// standing for row 32 b + 8 (2 h + i / 4) + 4 lk + i % 4; the membership operand uses the same
// row for each k slot, so every row of the tile enters each sum once. Lane (li, lk) of the
// result holds sequence 32 mb + 8 g + 4 lk + u (element 4 g + u) of channel c; the sums go to
// an LDS stage (FP32, row stride kStageStride), from which 16 threads per sequence finish it:
// a sequence that lies wholly in the tile gets its mean rounded to BF16 (16-byte stores) and
// the squared norm of that rounding (each thread's 8 channels in order, then a fixed xor
// tree), the one cut by the tile's start leaves its tail sums and the one cut by its end its
// head sums, which encode_finish_kernel joins. Every thread of the workgroup calls this with
// the same cnt.
__device__ __forceinline__ void enc_pool(const f32x16 (&acc)[4], const float* cap,
                                         const EncSeqs& seq, uint32_t cnt, uint32_t s0,
                                         uint32_t t, float* stage, const EncPool& p) {
    const uint32_t lane = enc_lane(), tid = 64 * enc_wave() + lane, li = lane & 31u, lk = lane >> 5;
    const uint32_t c = 32 * (tid >> 6) + li;
    bf16x8 hb[4][2];
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b)
#pragma unroll
        for (uint32_t g = 0; g < 4; ++g) {
            const float4 c4 = *reinterpret_cast<const float4*>(cap + 32 * b + 8 * g + 4 * lk);
            const float cu[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u)
                hb[b][g / 2][4 * (g % 2) + u] = f2bf(enc_relu(acc[b][4 * g + u], cu[u]));
        }
    for (uint32_t mb = 0; mb < (cnt + 31) / 32; ++mb) {
        const uint32_t jm = 32 * mb + li;
        const int32_t lo = jm < cnt ? seq.lo[jm] : 0;
        const int32_t hi = jm < cnt ? lo + int32_t(seq.len[jm]) : 0;
        // the sequence's rows [lo, hi) of the tile as a 128-bit mask, 32 rows per word
        const auto below = [](int32_t k) -> uint32_t {  // bits [0, k), k clamped to [0, 32]
            return k <= 0 ? 0u : k >= 32 ? 0xFFFFFFFFu : (1u << k) - 1u;
        };
        f32x16 ps = f32x16{0.f};
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t word = below(hi - int32_t(32 * b)) & ~below(lo - int32_t(32 * b));
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                // k slots 4 q + u stand for row 32 b + 16 h + 8 q + 4 lk + u: two slots per
                // dword, BF16 1.0 (0x3F80) where the row's bit is set
                const uint32_t x = word >> (16 * h + 4 * lk);
                u32x4 pk;
#pragma unroll
                for (uint32_t d = 0; d < 4; ++d) {
                    const uint32_t two = (x >> (8 * (d / 2) + 2 * (d % 2))) & 3u;
                    pk[d] = ((two * 0x8001u) & 0x10001u) * 0x3F80u;
                }
                ps = mfma(__builtin_bit_cast(bf16x8, pk), hb[b][h], ps);
            }
        }
#pragma unroll
        for (uint32_t e = 0; e < 16; ++e)
            stage[(8 * (e / 4) + 4 * lk + e % 4) * kStageStride + c] = ps[e];
        __syncthreads();
#pragma unroll
        for (uint32_t half = 0; half < 2; ++half) {
            const uint32_t idx = tid + 256 * half, jj = idx / 16, piece = idx % 16;
            const uint32_t j = 32 * mb + jj;
            const float4* src = reinterpret_cast<const float4*>(stage + jj * kStageStride + 8 * piece);
            const float4 x0 = src[0], x1 = src[1];
            const float sums[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            const bool in = j < cnt;
            const int32_t jlo = in ? seq.lo[j] : 0, jhi = in ? jlo + int32_t(seq.len[j]) : 0;
            const float inv = in ? seq.inv[j] : 0.f;
            bf16x8 v;
            float sq = 0.f;
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) {
                v[i] = f2bf(sums[i] * inv);
                sq += bf2f(v[i]) * bf2f(v[i]);
            }
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) sq += __shfl_xor(sq, o);
            if (in && jlo >= 0 && jhi <= int32_t(kEncRows)) {
                *reinterpret_cast<bf16x8*>(p.eb + uint64_t(s0 + j) * kChannels + 8 * piece) = v;
                if (piece == 0) p.norm[s0 + j] = sq;
            } else if (in && (jlo < 0 ? jhi > 0 : jhi > int32_t(kEncRows))) {
                float4* dst = reinterpret_cast<float4*>(
                    p.part + (uint64_t(t) * 2 + (jlo < 0 ? 1 : 0)) * kChannels + 8 * piece);
                dst[0] = x0;
                dst[1] = x1;
            }
        }
        __syncthreads();  // the stage is read before the next block of sequences overwrites it
    }
}

// The inputs of a tile land in the last kEncInBytes of regA by DMA (buffer loads into LDS, no
// registers), issued during the previous tile's conv3 while regA is free: the X0 window in the
// kXStride layout (kXSlots 16-byte slots, five per row, the fifth the row's padding), the valid
// bytes of rows r0 - 4 .. r0 + 159 (the three frames' masks at offsets 0, 2 and 4) and the
// offsets of the sequences s0 .. s0 + 95. Every X0 and valid read stays inside its buffer: the
// caller's buffers hold zero guard rows (8 X0 rows and 16 valid bytes before row 0, 8 X0 rows and
// 64 valid bytes past R), and padding slots and lanes past the window read a deliberately
// out-of-range offset, whose slot content is never used; the offsets past n are never used.
constexpr uint32_t kXSlots = 768;  // 12 per wave-instruction group: 4 waves x 3 x 64 lanes
constexpr uint32_t kInX0 = 0, kInValid = kXSlots * 16, kInOff = kInValid + 256;
constexpr uint32_t kEncInBytes = kInOff + 768;
constexpr uint32_t kEncIn = kEncRegionA - kEncInBytes;  // byte offset of the inputs in regA
static_assert(kXSlots >= kXRows * 5 && kEncIn % 16 == 0, "X0 window slots");
constexpr uint32_t kValidWindow = kEncMaskRows + 4;  // bytes: rows r0 - 4 .. r0 + 159
constexpr uint32_t kOutOfRange = 0xFFFFFFF0u;

struct EncSrc {
    __amdgpu_buffer_rsrc_t x0, valid, off;
};

__device__ __forceinline__ void enc_fetch(char* regA, uint32_t t, uint32_t s0,
                                          const EncSrc& src) {
    const uint32_t lane = enc_lane(), w = enc_wave();
    const int32_t r0 = int32_t(t * kEncRows);
    char* in = regA + kEncIn;
#pragma unroll
    for (uint32_t i = 0; i < 3; ++i) {
        const uint32_t q = (3 * w + i) * 64, slot = q + lane, row = slot / 5, piece = slot % 5;
        const uint32_t o = piece < 4 && row < kXRows
                               ? uint32_t(r0 + 2 + int32_t(row)) * (kEmbDim * 2) + 16 * piece
                               : kOutOfRange;  // window row r0 - 6 + row, 8 guard rows first
        __builtin_amdgcn_raw_ptr_buffer_load_lds(src.x0, (lds_void)(in + kInX0 + 16 * q), 16,
                                                 int(o), 0, 0, 0);
    }
    if (w == 0) {
        const uint32_t o =  // byte r0 - 4 + 4 lane, 16 guard bytes first
            4 * lane < kValidWindow ? uint32_t(r0 + 12) + 4 * lane : kOutOfRange;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(src.valid, (lds_void)(in + kInValid), 4, int(o),
                                                 0, 0, 0);
    } else {
        const uint32_t q = 64 * (w - 1);  // dwords q .. q + 63 of the offsets from s0
        __builtin_amdgcn_raw_ptr_buffer_load_lds(src.off, (lds_void)(in + kInOff + 4 * q), 4,
                                                 int(8 * s0 + 4 * (q + lane)), 0, 0, 0);
    }
}

// After the tile's DMA has landed (every wave waited for its own, then a barrier): turns the
// valid window in regA (which conv1's output overwrites) into the layers' ReLU caps (+inf on a
// valid row, 0 on any other; enc_relu) and builds the sequence table,
// returning the number of sequences meeting the tile (those from s0 on that start before its
// end; every thread gets the count after a barrier) and recording the one that runs past its
// end, if any.
__device__ __forceinline__ uint32_t enc_put(const char* regA, uint32_t t, uint32_t s0,
                                            uint64_t off0, float* cap, EncSeqs& seq,
                                            const EncPool& pool) {
    const uint32_t tid = threadIdx.x;
    const char* in = regA + kEncIn;
    if (tid < kValidWindow)
        cap[tid] = reinterpret_cast<const uint8_t*>(in + kInValid)[tid] ? __builtin_inff() : 0.f;
    bool meets = false;
    if (tid < kEncMaxSeqs && s0 + tid < pool.n) {
        const uint64_t* off = reinterpret_cast<const uint64_t*>(in + kInOff);
        const uint64_t o0 = off[tid], o1 = off[tid + 1];
        const uint32_t s = s0 + tid;
        const int64_t lo = int64_t(o0 - off0 + 2ull * s) - int64_t(t) * kEncRows;
        const uint32_t len = uint32_t(o1 - o0);
        meets = lo < int64_t(kEncRows);
        seq.lo[tid] = int32_t(lo);
        seq.len[tid] = len;
        seq.inv[tid] = len ? 1.f / float(len) : 0.f;
    }
    const uint32_t cnt = uint32_t(__syncthreads_count(meets));
    if (tid == 0) {  // the last sequence meeting the tile, when it runs past the tile's end
        const bool head = cnt && seq.lo[cnt - 1] >= 0 &&
                          seq.lo[cnt - 1] + int32_t(seq.len[cnt - 1]) > int32_t(kEncRows);
        pool.head_seq[t] = head ? s0 + cnt - 1 : kNoSeq;
    }
    return cnt;
}

// Persistent: workgroup b runs tiles b, b + gridDim.x, ...; each tile's inputs arrive by DMA
// during the previous tile's conv3 and pooling, and the weight ring runs on from one layer and
// tile into the next. X0 and `valid` have the guard rows enc_fetch reads.
__global__ __launch_bounds__(256, 2) void encode_kernel(const __bf16* __restrict__ X0,
                                                        const uint8_t* __restrict__ valid,
                                                        const __bf16* __restrict__ B,
                                                        const float* __restrict__ P, uint64_t R,
                                                        uint32_t tiles, EncPool pool) {
    __shared__ __attribute__((aligned(16))) char regA[kEncRegionA];
    __shared__ __attribute__((aligned(16))) char regB[kEncRegionB];
    __shared__ __attribute__((aligned(16))) float cap[kValidWindow];  // frames at 0, 2, 4
    __shared__ EncSeqs seq;
    // the layers' biases (LDS reads in the epilogues: a global load there would wait for the
    // weight ring's loads issued before it)
    __shared__ float bias[3][kChannels];
    uint32_t t = blockIdx.x;
    if (t >= tiles) return;  // the whole workgroup
    for (uint32_t e = threadIdx.x; e < 3 * kChannels; e += 256) {
        const uint32_t layer = e / kChannels, c = e % kChannels;
        bias[layer][c] = P[(layer == 0 ? kOffB1 : layer == 1 ? kOffB2 : kOffB3) + c];
    }
    const uint64_t off0 = pool.off[0];
    // the resources reach the guard rows: X0 from row -8 to R + 8, valid from byte -16 to R + 64
    const EncSrc src{
        __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(X0 - 8 * kEmbDim), 0,
                                          int((R + 16) * kEmbDim * 2), kBufferFlags),
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(valid - 16), 0, int(R + 80),
                                          kBufferFlags),
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(pool.off), 0,
                                          int((uint64_t(pool.n) + 1) * 8), kBufferFlags)};
    uint32_t s0 = pool.tile_first[t];
    enc_fetch(regA, t, s0, src);
    bf16x8 ring[kRingDepth];
    {
        const uint32_t lane = __lane_id();
        const __bf16* w1 = B + kOffWt1 + uint64_t(32 * (threadIdx.x >> 6) + (lane & 31u)) *
                                             (5 * kEmbDim) + 8 * (lane >> 5);
#pragma unroll
        for (uint32_t k = 0; k < kRingDepth; ++k)
            ring[k] = *reinterpret_cast<const bf16x8*>(w1 + 16 * k);
    }
    for (;;) {
        wait_vmcnt<0>();  // this wave's DMA of the tile's inputs (and the ring) has landed
        __syncthreads();
        const uint32_t cnt = enc_put(regA, t, s0, off0, cap, seq, pool);
        const uint32_t tn = t + gridDim.x;
        const bool more = tn < tiles;
        const uint32_t s0n = more ? pool.tile_first[tn] : 0;
        {
            f32x16 acc[5];
            enc_layer<kEmbDim, 5, false, kChannels>(regA + kEncIn + kInX0, kXStride, kH1Rows - 1,
                                                    B + kOffWt1, B + kOffWt2,
                                                    enc_bias_channels(bias[0]), ring, acc);
            __syncthreads();  // every wave is done reading X0 before H1 overwrites it
            enc_store<5, 1>(acc, cap, regA);
        }
        __syncthreads();
        {
            f32x16 acc[5];
            enc_layer<kChannels, 5, false, kChannels>(regA, kHStride, kH2Rows - 1, B + kOffWt2,
                                                      B + kOffWt3, enc_bias_channels(bias[1]),
                                                      ring, acc);
            enc_store<5, 1>(acc, cap + 2, regB);
        }
        __syncthreads();  // H2 is complete and every wave is done reading H1
        if (more) enc_fetch(regA, tn, s0n, src);
        {
            f32x16 acc[4];
            enc_layer<kChannels, 4, true, kEmbDim>(regB, kHStride, kEncRows - 1, B + kOffWt3,
                                                   B + kOffWt1, enc_bias_lane(bias[2]), ring,
                                                   acc);
            __syncthreads();  // every wave is done reading H2 before the pooling stage overwrites it
            enc_pool(acc, cap + 4, seq, cnt, s0, t, reinterpret_cast<float*>(regB), pool);
        }
        if (!more) break;
        t = tn;
        s0 = s0n;
    }
    // the ring's loads for a next tile that does not come land before the wave ends
    wait_vmcnt<0>();
}

// Joins the head sums of tile t and the tail sums of tile t + 1 of the sequence cut by their
// boundary. One workgroup of kChannels threads per tile.
__global__ __launch_bounds__(kChannels) void encode_finish_kernel(EncPool pool) {
    __shared__ float red[2][2];
    const uint32_t t = blockIdx.x, c = threadIdx.x;
    const uint32_t s = pool.head_seq[t];
    if (s == kNoSeq) return;  // the whole workgroup
    const float sum = pool.part[uint64_t(t) * 2 * kChannels + c] +
                      pool.part[(uint64_t(t + 1) * 2 + 1) * kChannels + c];
    const uint32_t len = uint32_t(pool.off[s + 1] - pool.off[s]);
    const float v = sum / float(len);
    const __bf16 vb = f2bf(v);
    pool.eb[uint64_t(s) * kChannels + c] = vb;
    float sq = bf2f(vb) * bf2f(vb);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    if ((c & 63u) == 0) red[0][c >> 6] = sq;
    __syncthreads();
    if (c == 0) pool.norm[s] = red[0][0] + red[0][1];
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
// Workgroup: 4 waves x 64 queries (four 16-row MFMA blocks per wave, operands kept in
// registers for the whole K = 128) against the references [r_begin, r_end) in tiles of 64,
// each consumed in quarters of 16 references by 16 v_mfma_f32_16x16x32_bf16 per wave.
// Reference tiles go from global memory straight into a four-buffer LDS ring by LDS-DMA loads
// (global_load_lds_dwordx4, no staging registers), issued three tiles ahead of the MFMAs that
// read them. A tile row of 128 BF16 (256 bytes) keeps its 16-byte pieces XOR-swizzled by the
// row (piece p of row r in slot p ^ (r % 16)): each wave's DMA writes one contiguous kilobyte
// and the operand reads of 16 consecutive rows hit 16 distinct bank groups.
//
// Pair (q, r) is kept when dot(q, r) - hq[q] >= hr[r], i.e. |q|^2 + |r|^2 - 2 dot <= tau, with
// hq = (|q|^2 - tau) / 2 and hr = |r|^2 / 2. Each MFMA chain starts from -hq held in registers
// that stay unchanged, and all accumulators of a lane belong to that lane's reference column,
// so a quarter without hits costs a max tree and one compare per column. Hits are extracted
// with one ballot per accumulator of the row groups that hold any, appended to the wave's LDS
// queue and flushed to the global arrays with one atomic per queue-full. The quarters run as a
// software pipeline inside each wave: the operands of quarter k + 1 are read from LDS and the
// hits of quarter k - 1 extracted while the MFMAs of quarter k run.
constexpr uint32_t kGramQ = 256;
constexpr uint32_t kTileR = 64;
constexpr uint32_t kRing = 4;
constexpr uint32_t kTilePieces = kTileR * kChannels / 8;  // 16-byte pieces per tile
constexpr uint32_t kLoadsPerTile = 5;  // DMA loads per wave and tile: 4 x 1 KB + the norms
constexpr uint32_t kNormSlot = 64;      // floats of a wave's norm DMA (16 rows, lanes 16+ repeat)
constexpr uint32_t kWaveQueue = 256;


// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The hit queues are written and read only through these asm statements, which the compiler
// does not see as memory operations: as visible LDS stores (which it cannot tell apart from the
// ring-slot DMA targets) they would make its wait-count pass drain all DMA in flight. Volatile
// asm statements keep their order, a wave's LDS operations execute in order, and the load waits
// for its own result before the asm ends.
__device__ __forceinline__ void queue_store(uint32_t* p, uint32_t v) {
    const uint32_t addr = uint32_t(reinterpret_cast<uintptr_t>((lds_void)p));
    asm volatile("ds_write_b32 %0, %1" ::"v"(addr), "v"(v));
}
__device__ __forceinline__ uint32_t queue_load(const uint32_t* p) {
    const uint32_t addr = uint32_t(reinterpret_cast<uintptr_t>((lds_void)(const_cast<uint32_t*>(p))));
    uint32_t v;
    asm volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(addr));
    return v;
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
            hit_q[g + i] = queue_load(qq + i);
            hit_r[g + i] = queue_load(qr + i);
        }
    wave_sync();
}

// One ring slot: a tile of 64 reference rows of 128 BF16, each row's 16-byte pieces
// XOR-swizzled (piece p of row r in slot p ^ (r % 16)), and their squared norms as four wave
// segments of kNormSlot floats (rows 16w .. 16w + 15 in segment w, entries 0 .. 15).
struct GramSlot {
    bf16x8 rows[kTilePieces];
    float norms[4 * kNormSlot];
};

// Per-lane byte offsets of the tile DMA (loop-invariant): wave w loads kilobytes 4w .. 4w + 3
// of a tile (rows 16w .. 16w + 15, pieces swizzled) and the norms of rows 16w .. 16w + 15
// (lanes 16 .. 63 repeat them, so that no lane is masked off).
struct GramDmaLanes {
    uint32_t rows[4];
    uint32_t norm;
};

__device__ __forceinline__ GramDmaLanes gram_dma_lanes() {
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    GramDmaLanes o;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
        const uint32_t row = 4 * (wave * 4 + u) + lane / 16;
        o.rows[u] = row * (kChannels * 2) + ((lane % 16) ^ (row % 16)) * 16;
    }
    o.norm = (wave * 16 + lane % 16) * 4;
    return o;
}

// DMA of tile t of the group (buffer resources over the group's rows and norms, so rows past
// the group's end read as zeros) into a slot. kLoadsPerTile buffer loads per wave, none of them
// under a branch, with loop-invariant lane offsets and the tile in the scalar offset.
__device__ __forceinline__ void gram_dma(GramSlot* dst, __amdgpu_buffer_rsrc_t rows,
                                         __amdgpu_buffer_rsrc_t norms, const GramDmaLanes& l,
                                         uint32_t t) {
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rows, (lds_void)(dst->rows + (wave * 4 + u) * 64),
                                                 16, l.rows[u], t * (kTileR * kChannels * 2), 0,
                                                 0);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(norms, (lds_void)(dst->norms + wave * kNormSlot), 4,
                                             l.norm, t * (kTileR * 4), 0, 0);
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// One quarter of a tile (16 reference rows) in registers: lane l holds row 16 qt + l % 16 of
// the slot, dims 32 ks + 8 (l / 16) .. + 7 in b[ks] (the second MFMA operand of k-step ks), and
// half that row's squared norm as its threshold (infinite past r_end).
struct GramQuarter {
    bf16x8 b[4];
    float hr;
};

// Per-lane byte offsets of the quarter reads in a slot: row 16 qt + n (n = l % 16), piece
// 4 ks + l / 16 at index piece ^ n, i.e. n * 256 + ((4 ks + l / 16) ^ n) * 16 + 4096 qt (16
// lanes of one piece column hit 16 distinct bank groups); the row's norm at
// norms[qt * kNormSlot + n].
struct GramReadLanes {
    uint32_t b[4];
    uint32_t norm;
};

__device__ __forceinline__ GramReadLanes gram_read_lanes() {
    const uint32_t lane = __lane_id(), n = lane & 15u, g = lane >> 4;
    GramReadLanes o;
#pragma unroll
    for (uint32_t ks = 0; ks < 4; ++ks) o.b[ks] = n * 256 + (((4 * ks + g) ^ n) * 16);
    o.norm = n * 4;
    return o;
}

// Quarter qt of a slot whose first row is reference `base`.
__device__ __forceinline__ void gram_read(const GramSlot* __restrict__ slot, uint32_t qt,
                                          const GramReadLanes& l, uint32_t base, uint32_t r_end,
                                          GramQuarter& g) {
    const char* rows = reinterpret_cast<const char*>(slot->rows) + 4096 * qt;
#pragma unroll
    for (uint32_t ks = 0; ks < 4; ++ks)
        g.b[ks] = *reinterpret_cast<const bf16x8*>(rows + l.b[ks]);
    const float norm = *reinterpret_cast<const float*>(
        reinterpret_cast<const char*>(slot->norms) + qt * kNormSlot * 4 + l.norm);
    g.hr = base + 16 * qt + (__lane_id() & 15u) < r_end ? 0.5f * norm : __builtin_inff();
}

// acc[qb] = -hq + dot over K = 128 of the wave's query block qb (16 queries) with the quarter's
// 16 references: lane l's element e is query 16 qb + 4 (l / 16) + e, reference l % 16.
__device__ __forceinline__ void gram_mfma(const bf16x8 (&a)[4][4], const f32x4 (&hneg)[4],
                                          const GramQuarter& g, f32x4 (&acc)[4]) {
#pragma unroll
    for (uint32_t ks = 0; ks < 4; ++ks)
#pragma unroll
        for (uint32_t qb = 0; qb < 4; ++qb)
            acc[qb] = mfma16(a[qb][ks], g.b[ks], ks ? acc[qb] : hneg[qb]);
}

struct GramHits {
    uint32_t* hit_q;
    uint32_t* hit_r;
    uint64_t cap;
    unsigned long long* count;
};

// The hits of a finished quarter, reference r in this lane: hr >= 0, and on non-negative floats
// the float order is the order of the bit patterns as signed integers, below which every
// negative float's pattern lies, so acc >= hr exactly when bits(acc) >= bits(hr) as int32. The
// maxima are taken on the bit patterns (integer max, no NaN canonicalisation; the accumulators
// hold no NaN). A quarter without hits costs the max tree and one compare.
__device__ __forceinline__ void gram_hits(const f32x4 (&acc)[4], float hr, uint32_t r,
                                          uint32_t q0, uint32_t* qq, uint32_t* qr, uint32_t& qn,
                                          const GramHits& out) {
    const uint32_t lane = __lane_id(), g4 = 4 * (lane >> 4);
    const int32_t hr_i = __float_as_int(hr);
    int32_t g[4];
#pragma unroll
    for (uint32_t qb = 0; qb < 4; ++qb)
        g[qb] = max(max(__float_as_int(acc[qb][0]), __float_as_int(acc[qb][1])),
                    max(__float_as_int(acc[qb][2]), __float_as_int(acc[qb][3])));
    const int32_t m = max(max(g[0], g[1]), max(g[2], g[3]));
    if (!__any(m >= hr_i)) return;
#pragma unroll
    for (uint32_t qb = 0; qb < 4; ++qb) {
        if (!__any(g[qb] >= hr_i)) continue;
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            const bool hit = acc[qb][e] >= hr;
            const uint64_t mask = __ballot(hit);
            if (!mask) continue;
            const uint32_t n = uint32_t(__popcll(mask));
            if (qn + n > kWaveQueue) {
                flush_queue(qq, qr, qn, out.hit_q, out.hit_r, out.cap, out.count);
                qn = 0;
            }
            if (hit) {
                const uint32_t pos =
                    qn + __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                                   __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
                queue_store(qq + pos, q0 + 16 * qb + g4 + e);
                queue_store(qr + pos, r);
            }
            qn += n;
        }
    }
}

// Wait until at most 2 x kLoadsPerTile of this wave's DMA are in flight, then a workgroup
// barrier. One asm statement, so that no DMA can be scheduled between the two.
__device__ __forceinline__ void gram_sync() {
    static_assert(kLoadsPerTile == 5 && kRing == 4, "the wait counts two tiles of 5 DMA");
    asm volatile("s_waitcnt vmcnt(10) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// The running state of the quarter pipeline: the quarter whose MFMAs come next (its operands
// read one quarter ahead) and the finished quarter whose hits are extracted after them, so
// each wave's LDS reads and hit tests overlap its own MFMAs.
struct GramRun {
    GramQuarter cur;
    f32x4 acc[4];  // the previous quarter's results
    float hr;      // and its lane's threshold and reference
    uint32_t r;
};

// Tile t (slot sl, quarter 0 already in run.cur) in four quarter steps; the last one first
// waits for tile t + 1 (slot sn), refills slot sl with tile t + 4 once every wave is done
// reading it, and reads tile t + 1's quarter 0.
__device__ __forceinline__ void gram_tile(GramSlot* __restrict__ sl, GramSlot* __restrict__ sn,
                                          uint32_t t, GramRun& run, const GramReadLanes& rl,
                                          const GramDmaLanes& dl, __amdgpu_buffer_rsrc_t rows,
                                          __amdgpu_buffer_rsrc_t norms, uint32_t r_begin,
                                          uint32_t r_end, const bf16x8 (&a)[4][4],
                                          const f32x4 (&hneg)[4], uint32_t q0, uint32_t* qq,
                                          uint32_t* qr, uint32_t& qn, const GramHits& out) {
    const uint32_t base = r_begin + t * kTileR, n16 = __lane_id() & 15u;
#pragma unroll
    for (uint32_t qt = 0; qt < 4; ++qt) {
        GramQuarter nxt;
        if (qt < 3) {
            gram_read(sl, qt + 1, rl, base, r_end, nxt);
        } else {
            gram_sync();
            gram_dma(sl, rows, norms, dl, t + 4);
            gram_read(sn, 0, rl, base + kTileR, r_end, nxt);
        }
        f32x4 acc[4];
        gram_mfma(a, hneg, run.cur, acc);
        gram_hits(run.acc, run.hr, run.r, q0, qq, qr, qn, out);
#pragma unroll
        for (uint32_t qb = 0; qb < 4; ++qb) run.acc[qb] = acc[qb];
        run.hr = run.cur.hr;
        run.r = base + 16 * qt + n16;
        run.cur = nxt;
    }
}

// The pipeline over tiles [0, ntiles4) of the group (ntiles4 a multiple of 4; tiles past the
// last real one hold no row before r_end and find no hits). Tile t sits in slot t % 4; the DMA
// of tile t + 4 goes into it once every wave has read tile t's last quarter. The four slots are
// separate restrict pointers, which tells the compiler (and its wait-count pass) that a DMA
// into one slot never aliases the reads of another, so the DMA of later tiles stays in flight
// over the current tile's MFMAs; its ordering is the waits in gram_sync.
__device__ __forceinline__ void gram_pipeline(GramSlot* __restrict__ s0,
                                              GramSlot* __restrict__ s1,
                                              GramSlot* __restrict__ s2,
                                              GramSlot* __restrict__ s3,
                                              uint32_t* __restrict__ qq,
                                              uint32_t* __restrict__ qr,
                                              __amdgpu_buffer_rsrc_t rows,
                                              __amdgpu_buffer_rsrc_t norms, uint32_t r_begin,
                                              uint32_t r_end, uint32_t ntiles4,
                                              const bf16x8 (&a)[4][4], const f32x4 (&hneg)[4],
                                              uint32_t q0, uint32_t& qn, const GramHits& out) {
    const GramDmaLanes dl = gram_dma_lanes();
    const GramReadLanes rl = gram_read_lanes();
    gram_dma(s0, rows, norms, dl, 0);
    gram_dma(s1, rows, norms, dl, 1);
    gram_dma(s2, rows, norms, dl, 2);
    gram_sync();  // tile 0 has landed
    gram_dma(s3, rows, norms, dl, 3);
    GramRun run;
    gram_read(s0, 0, rl, r_begin, r_end, run.cur);
#pragma unroll
    for (uint32_t qb = 0; qb < 4; ++qb) run.acc[qb] = f32x4{-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    run.hr = 0.f;
    run.r = 0;
#pragma unroll 1
    for (uint32_t t = 0; t < ntiles4; t += 4) {
        gram_tile(s0, s1, t, run, rl, dl, rows, norms, r_begin, r_end, a, hneg, q0, qq, qr, qn,
                  out);
        gram_tile(s1, s2, t + 1, run, rl, dl, rows, norms, r_begin, r_end, a, hneg, q0, qq, qr,
                  qn, out);
        gram_tile(s2, s3, t + 2, run, rl, dl, rows, norms, r_begin, r_end, a, hneg, q0, qq, qr,
                  qn, out);
        gram_tile(s3, s0, t + 3, run, rl, dl, rows, norms, r_begin, r_end, a, hneg, q0, qq, qr,
                  qn, out);
    }
    gram_hits(run.acc, run.hr, run.r, q0, qq, qr, qn, out);  // the last quarter
}

// Grid: q_blocks x groups workgroups in group-major order, query block fastest, renumbered
// per XCD (xcd_order): the workgroups of one reference group run together on the same XCD and
// share its tiles in that XCD's L2.
// The dispatcher hands workgroup b of a launch to XCD b % 8 (MI355X: 8 XCDs of 32 CUs, each
// with its own L2). Renumbering so that every XCD runs one contiguous range of the group-major
// order makes the workgroups resident on an XCD stream the same reference group through that
// XCD's L2 at the same time. A bijection on [0, n).
__device__ inline uint32_t xcd_order(uint32_t b, uint32_t n) {
    constexpr uint32_t kXcds = 8;
    const uint32_t x = b % kXcds, per = n / kXcds, rem = n % kXcds;
    return x * per + (x < rem ? x : rem) + b / kXcds;
}

__global__ __launch_bounds__(256, 2) void gram_kernel(const __bf16* __restrict__ Q,
                                                   const float* __restrict__ q_norm, uint32_t nq,
                                                   const __bf16* __restrict__ Rm,
                                                   const float* __restrict__ r_norm, uint32_t nr,
                                                   float tau, uint32_t r_per_group,
                                                   uint32_t q_blocks, uint32_t* hit_q,
                                                   uint32_t* hit_r, uint64_t cap,
                                                   unsigned long long* count) {
    // four separate objects, so that the compiler's alias analysis (and with it the wait-count
    // pass) tells a DMA into one slot from the reads of another
    __shared__ GramSlot slot0, slot1, slot2, slot3;
    __shared__ uint32_t queue_q[4][kWaveQueue], queue_r[4][kWaveQueue];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t id = xcd_order(blockIdx.x, gridDim.x);
    const uint32_t qb = id % q_blocks, grp = id / q_blocks;
    const uint32_t q0 = qb * kGramQ + wave * 64;
    const uint32_t r_begin = grp * r_per_group;
    const uint32_t r_end = r_begin + r_per_group < nr ? r_begin + r_per_group : nr;
    if (r_begin >= r_end) return;  // the whole workgroup, before any barrier
    const uint32_t ntiles = (r_end - r_begin + kTileR - 1) / kTileR;
    uint32_t qn = 0;  // wave-uniform queue fill (waves past the last query find no hits)
    const GramHits out{hit_q, hit_r, cap, count};

    // query block b (16 queries): lane l holds query q0 + 16 b + l % 16, dims
    // 32 ks + 8 (l / 16) .. + 7 in a[b][ks]; element e of its accumulator is query
    // q0 + 16 b + 4 (l / 16) + e, whose -hq starts the chain
    bf16x8 a[4][4];
    f32x4 hneg[4];
    // (every load unconditional, from a clamped row, so that none waits under a branch)
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
        const uint32_t q = q0 + 16 * b + (lane & 15u), qc = q < nq ? q : nq - 1;
#pragma unroll
        for (uint32_t ks = 0; ks < 4; ++ks) {
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(Q + uint64_t(qc) * kChannels +
                                                              ks * 32 + 8 * (lane >> 4));
            a[b][ks] = q < nq ? v : bf16x8{};
        }
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            const uint32_t qe = q0 + 16 * b + 4 * (lane >> 4) + e, qec = qe < nq ? qe : nq - 1;
            const float v = q_norm[qec];
            hneg[b][e] = qe < nq ? -0.5f * (v - tau) : -__builtin_inff();
        }
    }
    // the query operands are complete before the pipeline, whose waits count tile DMA only
    wait_vmcnt<0>();
    // buffer resources over the group's rows and norms (raw, bounds-checked: reads past the
    // group return zeros), the offsets in bytes from its first row
    const __amdgpu_buffer_rsrc_t rows = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<__bf16*>(Rm) + uint64_t(r_begin) * kChannels, 0,
        int((r_end - r_begin) * kChannels * 2), kBufferFlags);
    const __amdgpu_buffer_rsrc_t norms = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(r_norm) + r_begin, 0, int((r_end - r_begin) * 4), kBufferFlags);
    gram_pipeline(&slot0, &slot1, &slot2, &slot3, queue_q[wave], queue_r[wave], rows,
                  norms, r_begin, r_end, (ntiles + 3) / 4 * 4, a, hneg, q0, qn, out);
    if (qn) flush_queue(queue_q[wave], queue_r[wave], qn, hit_q, hit_r, cap, count);
    // no DMA may still be writing this workgroup's LDS when it ends
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
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
// Model distance |q|^2 + |r|^2 - 2 q.r of listed (query, reference) pairs: one wave per pair,
// two channels per lane, the dot product summed in FP32 across the wave.
__global__ __launch_bounds__(256) void pair_distance_kernel(
    const __bf16* __restrict__ Q, const float* __restrict__ qn, const __bf16* __restrict__ Rm,
    const float* __restrict__ rn, const uint32_t* __restrict__ pq, const uint32_t* __restrict__ pr,
    uint64_t n, float* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t t = (uint64_t(blockIdx.x) * 256 + threadIdx.x) >> 6;
    if (t >= n) return;  // whole waves leave together (t is uniform across a wave)
    const uint32_t a = pq[t], b = pr[t];
    const __bf16* x = Q + uint64_t(a) * kChannels + 2 * lane;
    const __bf16* y = Rm + uint64_t(b) * kChannels + 2 * lane;
    float dot = float(x[0]) * float(y[0]) + float(x[1]) * float(y[1]);
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off);
    if (lane == 0) out[t] = qn[a] + rn[b] - 2.0f * dot;
}

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
void launch_pair_distance(const __bf16* Q, const float* q_norm, const __bf16* Rm,
                          const float* r_norm, const uint32_t* pq, const uint32_t* pr, uint64_t n,
                          float* out, hipStream_t s) {
    if (n == 0) return;
    const uint64_t blocks = (n + 3) / 4;  // four waves (pairs) per workgroup
    hipLaunchKernelGGL(pair_distance_kernel, dim3(uint32_t(blocks)), dim3(256), 0, s, Q, q_norm,
                       Rm, r_norm, pq, pr, n, out);
}

void launch_embed(const uint8_t* codes, const uint64_t* off, uint32_t n, const __bf16* emb,
                  __bf16* X0, uint8_t* valid, uint32_t* tile_first, hipStream_t s) {
    const uint64_t total = uint64_t(n) * kEmbPieces;
    if (total == 0) return;
    hipLaunchKernelGGL(embed_kernel, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, s, codes,
                       off, n, emb, X0, valid, tile_first);
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

void launch_encode(const __bf16* X0, const uint8_t* valid, const __bf16* B, const float* P,
                   uint64_t R, const uint64_t* off, uint32_t n, const uint32_t* tile_first,
                   __bf16* eb, float* norm, float* part, uint32_t* head_seq, hipStream_t s) {
    if (R == 0) return;
    const EncPool pool{off, n, tile_first, eb, norm, part, head_seq};
    const uint32_t tiles = uint32_t(R / kEncRows);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const uint32_t grid = std::min<uint32_t>(tiles, 2u * uint32_t(cus));  // two per CU (LDS)
    hipLaunchKernelGGL(encode_kernel, dim3(grid), dim3(256), 0, s, X0, valid, B, P, R, tiles,
                       pool);
    hipLaunchKernelGGL(encode_finish_kernel, dim3(tiles), dim3(kChannels), 0, s, pool);
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
    static int per_cu = 0;  // resident workgroups per CU (the kernel's LDS and registers)
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void*>(gram_kernel), 256, 0) != hipSuccess || n < 1)
            n = 1;
        per_cu = n;
    }
    const uint64_t slots = uint64_t(per_cu) * cus;
    const uint32_t gq = (nq + kGramQ - 1) / kGramQ;
    const uint32_t tiles = (nr + kTileR - 1) / kTileR;
    // The workgroups run in rounds of `slots`; split the references into the number of groups
    // (whole tiles each, at least 16) that minimises rounds x (tiles per group + a few tiles of
    // pipeline fill).
    // (A group's rows are addressed by a buffer resource: fewer than 2^31 bytes each.)
    uint32_t per = std::min<uint32_t>(tiles, (1u << 31) / (kTileR * kChannels * 2) - 1);
    double best = -1.0;
    const uint32_t g_max = std::max<uint32_t>(1, std::min<uint32_t>(tiles / 16, 4096));
    for (uint32_t g = 1; g <= g_max; ++g) {
        const uint32_t pt = (tiles + g - 1) / g;
        const uint64_t wgs = uint64_t(gq) * ((tiles + pt - 1) / pt);
        if (g > 1 && wgs > 0x7FFFFFFFull) break;  // grid size limit
        if (uint64_t(pt) * kTileR * kChannels * 2 >= (1ull << 31)) continue;  // buffer range
        const double cost = double((wgs + slots - 1) / slots) * (pt + 4);
        if (best < 0 || cost < best) {
            best = cost;
            per = pt;
        }
    }
    const uint32_t groups = (tiles + per - 1) / per;
    hipLaunchKernelGGL(gram_kernel, dim3(gq * groups), dim3(256), 0, s, Q, q_norm, nq, Rm, r_norm,
                       nr, tau, per * kTileR, gq, hit_q, hit_r, cap, count);
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
