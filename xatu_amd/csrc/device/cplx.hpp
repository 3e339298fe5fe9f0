// Minimal complex arithmetic on HIP vector types (double2 / float2 = re, im).
#pragma once

#include <hip/hip_runtime.h>

namespace xatu_dev {

template <typename R>
struct Cx;
template <>
struct Cx<double> {
    using T = double2;
};
template <>
struct Cx<float> {
    using T = float2;
};
template <typename R>
using cx_t = typename Cx<R>::T;

__host__ __device__ inline double2 mk(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline float2 mk(float r, float i) { return make_float2(r, i); }

template <typename C>
__host__ __device__ inline C cadd(C a, C b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
template <typename C>
__host__ __device__ inline C csub(C a, C b) {
    a.x -= b.x;
    a.y -= b.y;
    return a;
}
template <typename C>
__host__ __device__ inline C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
// acc += a * b as two fused multiply-adds per component
__host__ __device__ inline void cfma(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
__host__ __device__ inline void cfma(float2& acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
}
// acc += conj(a) * b
__host__ __device__ inline void cfma_conj(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}
__host__ __device__ inline void cfma_conj(float2& acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(-a.y, b.x, acc.y);
}
template <typename C>
__host__ __device__ inline C cconj(C a) {
    a.y = -a.y;
    return a;
}

}  // namespace xatu_dev
