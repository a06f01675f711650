// Issue-rate probes: the measured integer-VALU peak that the scoring kernels' rooflines are
// quoted against, and the BF16 MFMA rate the prefilter's GEMMs are compared with.
//
// Each thread runs 16 independent chains of one VALU instruction, written as inline assembly
// so the opcode and operand kinds are exactly the ones named (registers only, no memory
// access inside the timed loop). 32 workgroups of 256 threads per CU keep every SIMD full.
// On gfx950 the result splits instructions into two issue classes: plain two-operand 32-bit
// and 16-bit ops on VGPRs issue a wave64 every 2 cycles per SIMD (128 lanes per CU per clock,
// the published FP32 vector rate), three-operand, packed, 64-bit and SGPR-operand forms every
// 4 cycles (64 lanes per CU per clock).

#include "st_internal.hpp"

namespace st {
namespace {

#define ST_PROBE_KERNEL(NAME, ASM)                                                         \
    __global__ __launch_bounds__(256) void NAME(uint32_t* out, uint32_t iters, uint32_t yv) { \
        uint32_t x[16];                                                                    \
        const uint32_t y = yv + threadIdx.x, z = yv ^ blockIdx.x;                          \
        _Pragma("unroll") for (int c = 0; c < 16; ++c) x[c] = threadIdx.x * 5u + c;        \
        for (uint32_t i = 0; i < iters; ++i) {                                             \
            _Pragma("unroll") for (int c = 0; c < 16; ++c)                                 \
                asm volatile(ASM : "+v"(x[c]) : "v"(y), "v"(z));                           \
        }                                                                                  \
        uint32_t s = 0;                                                                    \
        _Pragma("unroll") for (int c = 0; c < 16; ++c) s += x[c];                          \
        out[size_t(blockIdx.x) * 256 + threadIdx.x] = s;                                   \
    }

ST_PROBE_KERNEL(probe_add_u32, "v_add_u32 %0, %0, %1")
ST_PROBE_KERNEL(probe_max_i16, "v_max_i16 %0, %0, %1")
ST_PROBE_KERNEL(probe_max_i32, "v_max_i32 %0, %0, %1")
ST_PROBE_KERNEL(probe_pk_max_i16, "v_pk_max_i16 %0, %0, %1")
ST_PROBE_KERNEL(probe_max3_i32, "v_max3_i32 %0, %0, %1, %2")
ST_PROBE_KERNEL(probe_bfi_b32, "v_bfi_b32 %0, %0, %1, %2")

#undef ST_PROBE_KERNEL

// MFMA rate: four independent accumulator chains of v_mfma_f32_32x32x16_bf16 per wave.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void probe_mfma_32x32x16_bf16(float* out, uint32_t iters) {
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = __bf16(float(threadIdx.x + i) * 0.001f);
        b[i] = __bf16(float(i) * 0.002f);
    }
    f32x16 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = f32x16{0.f};
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int e = 0; e < 16; ++e) s += acc[c][e];
    out[size_t(blockIdx.x) * 256 + threadIdx.x] = s;
}

using ProbeKernel = void (*)(uint32_t*, uint32_t, uint32_t);
constexpr ProbeKernel kProbes[] = {probe_add_u32, probe_max_i16, probe_max_i32,
                                   probe_pk_max_i16, probe_max3_i32, probe_bfi_b32};

}  // namespace

const char* probe_name(int kind) {
    static const char* names[] = {"v_add_u32", "v_max_i16", "v_max_i32", "v_pk_max_i16",
                                  "v_max3_i32", "v_bfi_b32"};
    return kind >= 0 && kind < int(sizeof(names) / sizeof(names[0])) ? names[kind] : nullptr;
}

void launch_probe(int kind, uint32_t* out, uint32_t blocks, uint32_t iters, hipStream_t s) {
    hipLaunchKernelGGL(kProbes[kind], dim3(blocks), dim3(256), 0, s, out, iters, 12345u);
}

// LDS read rate: each thread runs 16 reads of W dwords per iteration (ds_read_b32 / _b64 /
// _b128 for W = 1, 2, 4). Read c of lane l takes W consecutive dwords at (c * 64 + l) * W,
// so a wave's 64 lanes cover 64 * W consecutive dwords, the bank-conflict-free pattern. An
// iteration's addresses carry an offset computed from the previous iteration's first sum
// and a mask the host passes as zero, so the reads cannot be hoisted out of the loop, while
// the 16 reads of one iteration share a base and issue back to back.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int W>
__global__ __launch_bounds__(256) void probe_lds_kernel(uint32_t* out, uint32_t iters,
                                                        uint32_t mask) {
    __shared__ uint32_t buf[16 * 64 * W];
    for (uint32_t e = threadIdx.x; e < 16u * 64u * W; e += 256) buf[e] = e * 2654435761u;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = lane + c;
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t* base = buf + lane * W + (acc[0] & mask);
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint32_t* at = base + c * 64 * W;
            if constexpr (W == 1) {
                acc[c] += *at;
            } else if constexpr (W == 2) {
                const u32x2 v = *reinterpret_cast<const u32x2*>(at);
                acc[c] += v.x + v.y;
            } else {
                const u32x4 v = *reinterpret_cast<const u32x4*>(at);
                acc[c] += (v.x + v.y) + (v.z + v.w);
            }
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int c = 0; c < 16; ++c) s += acc[c];
    out[size_t(blockIdx.x) * 256 + threadIdx.x] = s;
}

double launch_probe_lds(int dwords, uint32_t* out, uint32_t blocks, uint32_t iters,
                        hipStream_t s) {
    if (dwords == 4)
        hipLaunchKernelGGL(probe_lds_kernel<4>, dim3(blocks), dim3(256), 0, s, out, iters, 0u);
    else if (dwords == 2)
        hipLaunchKernelGGL(probe_lds_kernel<2>, dim3(blocks), dim3(256), 0, s, out, iters, 0u);
    else
        hipLaunchKernelGGL(probe_lds_kernel<1>, dim3(blocks), dim3(256), 0, s, out, iters, 0u);
    const int w = dwords == 4 || dwords == 2 ? dwords : 1;
    return double(blocks) * 256 * iters * 16 * w * 4;
}

// HBM write rate: a grid-stride fill of `n` 16-byte vectors (the store pattern of the
// write-bound distance kernels).
__global__ __launch_bounds__(256) void probe_fill_kernel(u32x4* __restrict__ p, uint64_t n,
                                                         uint32_t v) {
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    for (uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
        __builtin_nontemporal_store(u32x4{v, v + 1, v + 2, v + 3}, p + i);
}

void launch_probe_fill(void* p, uint64_t bytes, uint32_t blocks, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(probe_fill_kernel, dim3(blocks), dim3(256), 0, s,
                       static_cast<u32x4*>(p), bytes / 16, v);
}

double launch_probe_mfma(float* out, uint32_t blocks, uint32_t iters, hipStream_t s) {
    hipLaunchKernelGGL(probe_mfma_32x32x16_bf16, dim3(blocks), dim3(256), 0, s, out, iters);
    return double(blocks) * 4 /* waves */ * iters * 4 /* chains */ * 32.0 * 32 * 16 * 2;
}

}  // namespace st
