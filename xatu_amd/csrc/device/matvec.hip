// Dense products Y = H X with the resident BSE matrix.
#include "context.hpp"

namespace xatu_dev {

void launch_matvec64(const Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    (void)s;  // rocBLAS runs on the handle's stream, which is the context stream
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    const int n = static_cast<int>(c.N);
    XATU_BLAS(rocblas_zgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, nvec, n, &one,
                            reinterpret_cast<const rocblas_double_complex*>(c.d_H64.p), n,
                            reinterpret_cast<const rocblas_double_complex*>(X), n, &zero,
                            reinterpret_cast<rocblas_double_complex*>(Y), n));
}

void launch_matvec32(const Context& c, const float2* X, float2* Y, int nvec, hipStream_t s) {
    (void)s;
    const rocblas_float_complex one{1.0f, 0.0f}, zero{0.0f, 0.0f};
    const int n = static_cast<int>(c.N);
    XATU_BLAS(rocblas_cgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, nvec, n, &one,
                            reinterpret_cast<const rocblas_float_complex*>(c.d_H32.p), n,
                            reinterpret_cast<const rocblas_float_complex*>(X), n, &zero,
                            reinterpret_cast<rocblas_float_complex*>(Y), n));
}

void launch_matfree64(const Context&, const double2*, double2*, int, hipStream_t) {
    throw Error(XATU_ERR_INVALID_ARGUMENT, "matrix-free product not available yet");
}

void launch_matfree32(const Context&, const float2*, float2*, int, hipStream_t) {
    throw Error(XATU_ERR_INVALID_ARGUMENT, "matrix-free product not available yet");
}

}  // namespace xatu_dev
