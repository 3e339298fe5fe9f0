// gfx950 matrix-core helpers shared by the product kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace xatu_dev {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename R>
struct Mfma;
// 16x16x4 products. Both take A[i][k] from lane i + 16 k and B[k][j] from lane
// j + 16 k. Output register v of lane l holds C[row][l % 16] with row = 4 (l / 16) + v
// for the FP32 instruction and row = l / 16 + 4 v for the FP64 one (checked on gfx950).
template <>
struct Mfma<float> {
    using acc_t = f32x4;
    __device__ static acc_t mma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int out_row(int lk, int v) { return 4 * lk + v; }
};
template <>
struct Mfma<double> {
    using acc_t = f64x4;
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int out_row(int lk, int v) { return lk + 4 * v; }
};

// 16x16x32 FP16 product with FP32 accumulation (gfx950): lane l supplies 8 values
// A[l % 16][k], B[k][l % 16] for one group of 8 k indices selected by l / 16 (the same
// group for A and B); the output has the FP32 16x16x4 layout, row = 4 (l / 16) + v
// (checked on gfx950).
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
__device__ inline f32x4 mma_f16(f16x8 a, f16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// Vectors per product launch (two 16-column MFMA tiles).
constexpr int kBlockVectors = 32;

// Rows per workgroup of the split product; with the FP32 matrix built block-lower, column i
// holds rows k >= kSplitRowBlock floor(i / kSplitRowBlock) only.
constexpr int kSplitRowBlock = 256;

// Number of XCDs of the MI355X; consecutive workgroup ids are dealt round-robin to them.
constexpr int kXcds = 8;

}  // namespace xatu_dev
