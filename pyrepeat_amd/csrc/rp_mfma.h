// MFMA fragment layouts on gfx950 (wave64), checked on the device.
#pragma once
#include <hip/hip_runtime.h>

namespace rp {

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec16 __attribute__((ext_vector_type(16)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef int ivec16 __attribute__((ext_vector_type(16)));
typedef int ivec4 __attribute__((ext_vector_type(4)));
typedef _Float16 hvec8 __attribute__((ext_vector_type(8)));

// v_mfma_f64_16x16x4f64:
//   A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
//   D[v] of lane l is element (row l/16 + 4v, col l%16).
struct F64x16 {
    static __device__ __forceinline__ int a_row(int lane) { return lane & 15; }
    static __device__ __forceinline__ int a_k(int lane) { return lane >> 4; }
    static __device__ __forceinline__ int d_row(int lane, int v) { return (lane >> 4) + 4 * v; }
    static __device__ __forceinline__ int d_col(int lane) { return lane & 15; }
    static __device__ __forceinline__ dvec4 mma(double a, double b, dvec4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// v_mfma_f32_32x32x16_f16 / v_mfma_i32_32x32x32_i8:
//   A[i][k] lane i%32 + 32*(k / kpl), element k % kpl (kpl = 8 for f16, 16 for i8),
//   D[r] of lane l is element (row 8*(r/4) + 4*(l/32) + r%4, col l%32).
struct M32 {
    static __device__ __forceinline__ int d_row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
    static __device__ __forceinline__ int d_col(int lane) { return lane & 31; }
};

}  // namespace rp
