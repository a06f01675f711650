// VALU issue-rate probe: the measured integer-VALU peak that the scoring kernels' rooflines
// are quoted against.
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

}  // namespace st
