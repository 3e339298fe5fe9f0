// Element-wise kernels used by the Davidson solver.
#include "context.hpp"
#include "cplx.hpp"

namespace xatu_dev {

namespace {

// R[:, j] = AX[:, j] - theta_j X[:, j]; norms2[j] += |R[:, j]|^2 via block partials.
// grid: (ceil(N/256), nvec)
__global__ void residual_kernel(const double2* __restrict__ X, const double2* __restrict__ AX,
                                const double* __restrict__ theta, int64_t N, double2* __restrict__ R,
                                double* __restrict__ partial) {
    __shared__ double s[256];
    const int j = blockIdx.y;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < N) {
        const double2 x = X[i + j * N];
        const double2 ax = AX[i + j * N];
        const double th = theta[j];
        const double2 r = make_double2(ax.x - th * x.x, ax.y - th * x.y);
        R[i + j * N] = r;
        acc = r.x * r.x + r.y * r.y;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[static_cast<size_t>(j) * gridDim.x + blockIdx.x] = s[0];
}

// Diagonal (Davidson) preconditioner: T[:, j] = R[:, j] / (theta_j - d), with the
// denominator kept away from zero.
__global__ void precondition_kernel(const double2* __restrict__ R, const double* __restrict__ theta,
                                    const double* __restrict__ diag, int64_t N, double2* __restrict__ T) {
    const int j = blockIdx.y;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double den = theta[j] - diag[i];
    if (fabs(den) < 1e-3) den = (den >= 0.0) ? 1e-3 : -1e-3;
    const double2 r = R[i + j * N];
    T[i + j * N] = make_double2(r.x / den, r.y / den);
}

// dst[i] = src[i] for n complex entries (non-overlapping), 2 per thread
__global__ void copy_vectors_kernel(const double2* __restrict__ src, int64_t n, double2* __restrict__ dst) {
    const int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 2;
    if (i < n) dst[i] = src[i];
    if (i + 1 < n) dst[i + 1] = src[i + 1];
}

// dst = [a (na), b (nb), c (nc)]: three device arrays gathered into one destination (mapped
// host memory), one double per thread
__global__ void gather3_kernel(const double* __restrict__ a, int na, const double* __restrict__ b, int nb,
                               const double* __restrict__ c, int nc, double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) {
        dst[i] = a[i];
    } else if (i < na + nb) {
        dst[i] = b[i - na];
    } else if (i < na + nb + nc) {
        dst[i] = c[i - na - nb];
    }
}

// V[:, j] = e_{idx[j]}
__global__ void unit_vectors_kernel(const int* __restrict__ idx, int64_t N, int nvec, double2* __restrict__ V) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nvec) V[idx[j] + static_cast<int64_t>(j) * N] = make_double2(1.0, 0.0);
}

// out[j + 3 n] = sum_i X[i + n N] vel[i + j N] (no conjugation, as the reference's
// oscillator sum): one workgroup per (j, n), fixed-order tree reduction.
__global__ __launch_bounds__(256) void oscillator_kernel(const double2* __restrict__ X, int64_t N,
                                                         const double2* __restrict__ vel, double2* __restrict__ out) {
    __shared__ double2 s[256];
    const int j = blockIdx.x % 3, n = blockIdx.x / 3;
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t i = threadIdx.x; i < N; i += 256) cfma(acc, X[i + n * N], vel[i + j * N]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] = cadd(s[threadIdx.x], s[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

}  // namespace

void launch_oscillator(const double2* X, int64_t N, int nstates, const double2* vel, double2* out, hipStream_t s) {
    hipLaunchKernelGGL(oscillator_kernel, dim3(3 * nstates), dim3(256), 0, s, X, N, vel, out);
    XATU_HIP(hipGetLastError());
}

void launch_residual(const double2* X, const double2* AX, const double* theta, int64_t N, int nvec, double2* R,
                     double* partial, int* nblocks, hipStream_t s) {
    const int nbk = static_cast<int>((N + 255) / 256);
    *nblocks = nbk;
    hipLaunchKernelGGL(residual_kernel, dim3(nbk, nvec), dim3(256), 0, s, X, AX, theta, N, R, partial);
    XATU_HIP(hipGetLastError());
}

void launch_precondition(const double2* R, const double* theta, const double* diag, int64_t N, int nvec, double2* T,
                         hipStream_t s) {
    hipLaunchKernelGGL(precondition_kernel, dim3((N + 255) / 256, nvec), dim3(256), 0, s, R, theta, diag, N, T);
    XATU_HIP(hipGetLastError());
}

// Device-to-device copy of vector blocks as a kernel on the stream (no copy-engine
// hand-over between the solver's kernels).
void launch_copy_vectors(const double2* src, int64_t n, double2* dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(copy_vectors_kernel, dim3(static_cast<unsigned>((n + 511) / 512)), dim3(256), 0, s, src, n, dst);
    XATU_HIP(hipGetLastError());
}

void launch_gather3(const double* a, int na, const double* b, int nb, const double* c, int nc, double* dst,
                    hipStream_t s) {
    const int n = na + nb + nc;
    if (n <= 0) return;
    hipLaunchKernelGGL(gather3_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, na, b, nb, c, nc, dst);
    XATU_HIP(hipGetLastError());
}

void launch_unit_vectors(const int* idx, int64_t N, int nvec, double2* V, hipStream_t s) {
    hipLaunchKernelGGL(unit_vectors_kernel, dim3((nvec + 63) / 64), dim3(64), 0, s, idx, N, nvec, V);
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
