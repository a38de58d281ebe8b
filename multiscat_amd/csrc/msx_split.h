// Helpers of the split-FP16 kernels (msx_diag_h3.hip, msx_dft_h3.hip): an FP32 operand a is
// carried as a pair of FP16 values under a power-of-two scale s, s a = hi + lo with
// hi = fp16(s a) and lo = fp16(s a - hi) (22 significant bits), for the 16x16x32 FP16 MFMA.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace msx {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// hi and lo of two elements in the low and high halves of one register each: every half is one
// mixed-precision FMA that rounds once to FP16 (s is a power of two, so s v - hi is exact before
// that rounding). The compiler, left to itself, formed a multiply, a conversion back and a
// subtraction for some of them.
__device__ __forceinline__ uint32_t split_hi2(float v0, float v1, float s) {
    uint32_t r;
    asm("v_fma_mixlo_f16 %0, %1, %2, 0\n\tv_fma_mixhi_f16 %0, %3, %2, 0"
        : "=&v"(r) : "v"(v0), "v"(s), "v"(v1));
    return r;
}
__device__ __forceinline__ uint32_t split_lo2(float v0, float v1, float s, uint32_t hi) {
    uint32_t r;
    asm("v_fma_mixlo_f16 %0, %1, %2, -%4 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %0, %3, %2, -%4 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
        : "=&v"(r) : "v"(v0), "v"(s), "v"(v1), "v"(hi));
    return r;
}
// eight elements (s v[u], u = 0..7) into one hi and one lo operand register quadruple
__device__ __forceinline__ void split8s(const float (&v)[8], float s, h8 &hi, h8 &lo) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    u4 h, l;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        h[q] = split_hi2(v[2 * q], v[2 * q + 1], s);
        l[q] = split_lo2(v[2 * q], v[2 * q + 1], s, h[q]);
    }
    hi = __builtin_bit_cast(h8, h);
    lo = __builtin_bit_cast(h8, l);
}

// max(|a|, |b|, m) in one instruction (the abs modifiers are free; fmaxf under the IEEE mode
// of compute kernels quiets every input first, two extra instructions per element)
__device__ __forceinline__ float max3abs(float a, float b, float m) {
    float r;
    asm("v_max3_f32 %0, |%1|, |%2|, %3" : "=v"(r) : "v"(a), "v"(b), "v"(m));
    return r;
}

// Power-of-two exponent e with 2^e m in [2^(top-1), 2^top) for a maximum m >= 0 (m = 0 or tiny:
// at most top + 100, which keeps the scaled values finite).
__device__ __forceinline__ int scale_exp(float m, int top) {
    int e2 = 0;
    (void)frexpf(m, &e2);  // m = f 2^e2, f in [0.5, 1)
    return top - max(e2, -100);
}

// element at a 32-bit byte offset from a uniform base: one VGPR offset on the scalar base
template <typename T>
__device__ __forceinline__ const T &at(const T *base, uint32_t byte_off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ T &at(T *base, uint32_t byte_off) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte_off);
}

}  // namespace msx
