// 16x16x4 matrix-core (MFMA) wrapper shared by the GEMM-shaped kernels, FP32 and FP64.
//
// Operand layout of one 16x16x4 step on a 64-lane wavefront (g = lane / 16, l = lane % 16):
//   A[i][k]  (16 x 4)  lane i + 16 k
//   B[k][j]  (4 x 16)  lane j + 16 k
//   C[i][j]  (16 x 16) lane j + 16 g, accumulator register v, with
//            i = 4 g + v for the FP32 instruction and i = 4 v + g for the FP64 one
//            (measured on gfx950: row(g, v) below).
#pragma once

#include <hip/hip_runtime.h>

namespace msx {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    using acc = v4f;
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // row inside the 16 x 16 tile of accumulator register v in lane group g
    static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
    // Source row to feed as A row i so that register v of lane group g holds source row 4 g + v
    // (perm(row(g, v)) == 4 g + v).
    static __device__ __forceinline__ int perm(int i) { return i; }
};
template <>
struct Mfma<double> {
    using acc = v4d;
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int v) { return 4 * v + g; }
    static __device__ __forceinline__ int perm(int i) { return 4 * (i & 3) + (i >> 2); }
};

// Row stride (in elements) >= n with stride % 64 == 36: the 16-row x 4-column A-fragment read
// of a wavefront then touches 64 distinct 4-byte LDS banks.
__host__ __device__ constexpr int lds_stride_a(int n) { return ((n + 27) / 64) * 64 + 36; }

// Row stride >= n with stride % 64 == 16: the 4-row x 16-column B-fragment read then touches
// 64 distinct banks (FP32) or 32 distinct bank pairs per half-wavefront (FP64).
__host__ __device__ constexpr int lds_stride_b(int n) { return ((n + 47) / 64) * 64 + 16; }

}  // namespace msx
