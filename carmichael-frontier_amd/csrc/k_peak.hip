// k_peak.hip -- this GPU's vector-ALU issue rate for the integer instructions the join's kernels
// are built from (the denominator of the integer-ALU roofline the benchmark reports).
//
// Every thread runs `iters` rounds of CHAINS independent chains of one instruction, so each SIMD
// always has ready instructions; the instructions are written in inline asm on register operands
// only (no memory operand), so the compiler can neither fold nor reorder them away. The grid puts
// several waves on every SIMD of every CU. The rate is wave-level VALU instructions per second (one
// wave instruction is 64 lane operations).
#include "cf_internal.h"

namespace {
constexpr int PK_THREADS = 256;
constexpr int CHAINS = 8;
constexpr int UNROLL = 16;   // rounds of the CHAINS instructions per loop trip

template <int OP>
__device__ __forceinline__ void step(u32 &a, u64 &w, u32 b) {
    if constexpr (OP == CF_PEAK_ADD_U32) {
        asm volatile("v_add_u32 %0, %0, %1" : "+v"(a) : "v"(b));
    } else if constexpr (OP == CF_PEAK_MUL_LO_U32) {
        asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a) : "v"(b));
    } else if constexpr (OP == CF_PEAK_MUL_HI_U32) {
        asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a) : "v"(b));
    } else {   // CF_PEAK_MAD_U64_U32: w = a * b + w, 64-bit accumulator
        asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(w) : "v"(a), "v"(b) : "vcc");
    }
}

template <int OP>
__global__ void __launch_bounds__(PK_THREADS) k_valu_peak(u32 *__restrict__ out, u32 iters, u32 seed) {
    u32 a[CHAINS];
    u64 w[CHAINS];
    const u32 b = seed | 1u;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) {
        a[c] = seed + threadIdx.x * 8u + (u32)c;
        w[c] = a[c];
    }
    for (u32 i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
            for (int c = 0; c < CHAINS; ++c) step<OP>(a[c], w[c], b);
        }
    }
    u32 x = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x ^= a[c] ^ (u32)w[c] ^ (u32)(w[c] >> 32);
    out[(size_t)blockIdx.x * PK_THREADS + threadIdx.x] = x;
}
}  // namespace

extern "C" int cf_peak_valu(cf_ctx *ctx, int op, uint64_t iters, uint64_t *wave_instructions, double *ms) {
    API_BEGIN
    CF_REQUIRE(ctx && wave_instructions && ms, "NULL argument");
    CF_REQUIRE(op >= CF_PEAK_ADD_U32 && op <= CF_PEAK_MAD_U64_U32, "unknown instruction");
    CF_REQUIRE(iters >= 1 && iters <= (1u << 24), "need 1 <= iters <= 2^24");
    hipDeviceProp_t prop;
    CF_HIP(hipGetDeviceProperties(&prop, ctx->device));
    // eight 4-wave blocks per CU: eight waves on each of its four SIMDs
    const unsigned blocks = (unsigned)prop.multiProcessorCount * 8u;
    DBuf<u32> out(ctx, (size_t)blocks * PK_THREADS);
    auto launch = [&](u32 it) {
        switch (op) {
            case CF_PEAK_ADD_U32:
                CF_LAUNCH(ctx, "k_valu_peak", k_valu_peak<CF_PEAK_ADD_U32>, dim3(blocks), dim3(PK_THREADS), 0, out.p, it, 7u);
                break;
            case CF_PEAK_MUL_LO_U32:
                CF_LAUNCH(ctx, "k_valu_peak", k_valu_peak<CF_PEAK_MUL_LO_U32>, dim3(blocks), dim3(PK_THREADS), 0, out.p, it, 7u);
                break;
            case CF_PEAK_MUL_HI_U32:
                CF_LAUNCH(ctx, "k_valu_peak", k_valu_peak<CF_PEAK_MUL_HI_U32>, dim3(blocks), dim3(PK_THREADS), 0, out.p, it, 7u);
                break;
            default:
                CF_LAUNCH(ctx, "k_valu_peak", k_valu_peak<CF_PEAK_MAD_U64_U32>, dim3(blocks), dim3(PK_THREADS), 0, out.p, it, 7u);
        }
    };
    launch(1);   // code object loaded, clocks up
    hipEvent_t e0, e1;
    CF_HIP(hipEventCreate(&e0));
    CF_HIP(hipEventCreate(&e1));
    CF_HIP(hipEventRecord(e0, ctx->stream));
    launch((u32)iters);
    CF_HIP(hipEventRecord(e1, ctx->stream));
    CF_HIP(hipEventSynchronize(e1));
    float t = 0;
    CF_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t;
    *wave_instructions = (u64)blocks * (PK_THREADS / 64) * iters * UNROLL * CHAINS;
    API_END
}
