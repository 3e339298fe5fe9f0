// Dense products Y = H X with the resident BSE matrix on the matrix cores, and the
// split-K Gram kernel C = A^H B used by the Davidson projections.
//
// Product: H is column-major complex (FP32: float2, FP64: double2), N x N. The
// vectors are staged transposed, Xt[j][n] (32 vectors, zero padded). Each wave owns
// 64 rows (four 16-row tiles) and all 32 vectors (two 16-column tiles) and walks the
// columns of H four at a time: lane l loads the complex H(row = l%16 + 16 t, col =
// j + l/16), so the 16x16x4 MFMA sees A[i][k] = Re H(i, j+k) in one product and
// Im H(i, j+k) in the other. With B[k][n] built from X(j+k, n) the complex product
// is four real MFMA products per tile pair (Re Re, Im (-Im), Re Im, Im Re), exact in
// the sense of the usual 4-real-product complex multiply. Column ranges are split
// across workgroups (split-K) to fill all 256 CUs; partial results go to a workspace
// and a reduction kernel sums them and writes Y in the caller's column-major layout.
#include "context.hpp"
#include "cplx.hpp"

namespace xatu_dev {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kVec = 32;         // vectors per product launch
constexpr int kWaveRows = 64;    // rows per wave
constexpr int kBlockRows = 256;  // rows per workgroup (4 waves)

template <typename R>
struct Mfma;
// Both take A[i][k] from lane i + 16 k and B[k][j] from lane j + 16 k. Output
// register v of lane l holds C[row][l % 16] with row = 4 (l / 16) + v for the FP32
// instruction and row = l / 16 + 4 v for the FP64 one (checked on gfx950).
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

// Xt[j][n] = X[j + n N] for n < nvec, 0 otherwise (j < N, n < 32).
template <typename Rin, typename R>
__global__ void stage_vectors_kernel(const cx_t<Rin>* __restrict__ X, int64_t N, int nvec, cx_t<R>* __restrict__ Xt) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= N * kVec) return;
    const int64_t j = idx / kVec;
    const int n = static_cast<int>(idx % kVec);
    cx_t<R> v = mk(R(0), R(0));
    if (n < nvec) {
        const cx_t<Rin> x = X[j + n * N];
        v = mk(static_cast<R>(x.x), static_cast<R>(x.y));
    }
    Xt[idx] = v;
}

// grid: (ceil(N / 256), splits); block 256. Partial sums to Yp[split][row][32].
template <typename R>
__global__ __launch_bounds__(256) void dense_product_kernel(const cx_t<R>* __restrict__ H, const cx_t<R>* __restrict__ Xt,
                                                            int64_t N, int64_t kchunk, cx_t<R>* __restrict__ Yp) {
    using M = Mfma<R>;
    using acc_t = typename M::acc_t;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * kBlockRows + wave * kWaveRows;
    const int64_t kbeg = static_cast<int64_t>(blockIdx.y) * kchunk;
    const int64_t kend = min(N, kbeg + kchunk);
    const int li = lane & 15;  // row within a 16-row tile / vector within a 16-vector tile
    const int lk = lane >> 4;  // column offset 0..3 within the 4-column step

    acc_t accR[4][2], accI[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            accR[t][u] = acc_t{0, 0, 0, 0};
            accI[t][u] = acc_t{0, 0, 0, 0};
        }

    int64_t rows[4];
    bool rok[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        rows[t] = row0 + 16 * t + li;
        rok[t] = rows[t] < N;
    }

    for (int64_t j = kbeg; j < kend; j += 4) {
        const int64_t col = j + lk;
        const bool cok = col < kend;
        cx_t<R> h[4], x[2];
#pragma unroll
        for (int t = 0; t < 4; ++t) h[t] = (cok && rok[t]) ? H[rows[t] + col * N] : mk(R(0), R(0));
#pragma unroll
        for (int u = 0; u < 2; ++u) x[u] = cok ? Xt[col * kVec + 16 * u + li] : mk(R(0), R(0));
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                accR[t][u] = M::mma(h[t].x, x[u].x, accR[t][u]);
                accR[t][u] = M::mma(h[t].y, -x[u].y, accR[t][u]);
                accI[t][u] = M::mma(h[t].x, x[u].y, accI[t][u]);
                accI[t][u] = M::mma(h[t].y, x[u].x, accI[t][u]);
            }
    }
    cx_t<R>* out = Yp + static_cast<int64_t>(blockIdx.y) * N * kVec;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = row0 + 16 * t + M::out_row(lk, v);
            if (r < N) {
#pragma unroll
                for (int u = 0; u < 2; ++u) out[r * kVec + 16 * u + li] = mk(accR[t][u][v], accI[t][u][v]);
            }
        }
}

// Y[r + n N] = sum_s Yp[s][r][n] for n < nvec (written in the output precision).
template <typename R, typename Rout>
__global__ void reduce_product_kernel(const cx_t<R>* __restrict__ Yp, int64_t N, int nvec, int splits,
                                      cx_t<Rout>* __restrict__ Y) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (idx >= N || n >= nvec) return;
    Rout sr = 0, si = 0;
    for (int s = 0; s < splits; ++s) {
        const cx_t<R> v = Yp[(static_cast<int64_t>(s) * N + idx) * kVec + n];
        sr += static_cast<Rout>(v.x);
        si += static_cast<Rout>(v.y);
    }
    Y[idx + static_cast<int64_t>(n) * N] = mk(sr, si);
}

// ---------------------------------------------------------------------------
// Gram kernel: C (m x n) = A^H B over N rows, FP64 complex. Tile 64 x 32 per
// workgroup, each thread a 4 x 2 micro-tile; rows split into chunks, partial
// results reduced by a second kernel in a fixed order.
constexpr int kGM = 64, kGN = 32, kGR = 16;

__global__ __launch_bounds__(256) void gram_partial_kernel(const double2* __restrict__ A, int lda,
                                                           const double2* __restrict__ B, int ldb, int64_t N, int m,
                                                           int n, int64_t rows_per_chunk, double2* __restrict__ P) {
    __shared__ double2 sA[kGR][kGM + 1];
    __shared__ double2 sB[kGR][kGN + 1];
    const int tm = threadIdx.x % 16, tn = threadIdx.x / 16;
    const int m0 = blockIdx.x * kGM, n0 = blockIdx.y * kGN;
    const int64_t rbeg = static_cast<int64_t>(blockIdx.z) * rows_per_chunk;
    const int64_t rend = min(N, rbeg + rows_per_chunk);
    double2 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mk(0.0, 0.0);
    for (int64_t r0 = rbeg; r0 < rend; r0 += kGR) {
        for (int e = threadIdx.x; e < kGR * kGM; e += 256) {
            const int rr = e % kGR, cc = e / kGR;
            const int64_t r = r0 + rr;
            sA[rr][cc] = (r < rend && m0 + cc < m) ? A[r + static_cast<int64_t>(m0 + cc) * lda] : mk(0.0, 0.0);
        }
        for (int e = threadIdx.x; e < kGR * kGN; e += 256) {
            const int rr = e % kGR, cc = e / kGR;
            const int64_t r = r0 + rr;
            sB[rr][cc] = (r < rend && n0 + cc < n) ? B[r + static_cast<int64_t>(n0 + cc) * ldb] : mk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < kGR; ++rr) {
            double2 a[4], b[2];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[rr][tm + 16 * i];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = sB[rr][tn + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) cfma_conj(acc[i][j], a[i], b[j]);
        }
        __syncthreads();
    }
    double2* out = P + static_cast<int64_t>(blockIdx.z) * m * n;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int mi = m0 + tm + 16 * i, nj = n0 + tn + 16 * j;
            if (mi < m && nj < n) out[mi + static_cast<int64_t>(nj) * m] = acc[i][j];
        }
}

__global__ void gram_reduce_kernel(const double2* __restrict__ P, int m, int n, int chunks, double2* __restrict__ C,
                                   int ldc) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    const int i = idx % m, j = idx / m;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < chunks; ++c) {
        const double2 v = P[static_cast<int64_t>(c) * m * n + idx];
        sr += v.x;
        si += v.y;
    }
    C[i + static_cast<int64_t>(j) * ldc] = mk(sr, si);
}

template <typename Rin, typename R, typename Rout>
void dense_product(const Context& c, const cx_t<R>* H, const cx_t<Rin>* X, cx_t<Rout>* Y, int nvec, hipStream_t s,
                   DevBuf<cx_t<R>>& xt, DevBuf<cx_t<R>>& yp) {
    const int64_t N = c.N;
    const int row_blocks = static_cast<int>((N + kBlockRows - 1) / kBlockRows);
    int splits = std::max(1, std::min(64, (1024 + row_blocks - 1) / row_blocks));
    int64_t kchunk = ((N + splits - 1) / splits + 3) / 4 * 4;
    splits = static_cast<int>((N + kchunk - 1) / kchunk);
    xt.alloc(std::max<size_t>(xt.n, static_cast<size_t>(N) * kVec));
    yp.alloc(std::max<size_t>(yp.n, static_cast<size_t>(splits) * N * kVec));
    for (int v0 = 0; v0 < nvec; v0 += kVec) {
        const int nv = std::min(kVec, nvec - v0);
        const int64_t tot = N * kVec;
        hipLaunchKernelGGL((stage_vectors_kernel<Rin, R>), dim3((tot + 255) / 256), dim3(256), 0, s, X + v0 * N, N, nv,
                           xt.p);
        hipLaunchKernelGGL((dense_product_kernel<R>), dim3(row_blocks, splits), dim3(256), 0, s, H, xt.p, N, kchunk,
                           yp.p);
        hipLaunchKernelGGL((reduce_product_kernel<R, Rout>), dim3((N + 255) / 256, nv), dim3(256), 0, s, yp.p, N, nv,
                           splits, Y + v0 * N);
    }
    XATU_HIP(hipGetLastError());
}

}  // namespace

void launch_matvec64(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    dense_product<double, double, double>(c, c.d_H64.p, X, Y, nvec, s, c.s_xt64, c.s_yp64);
}

// FP32 product of FP64 vectors: staged in FP32, accumulated in FP32, widened on output.
void launch_matvec32(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    dense_product<double, float, double>(c, c.d_H32.p, X, Y, nvec, s, c.s_xt32, c.s_yp32);
}

void launch_gram(Context& c, const double2* A, int lda, const double2* B, int ldb, int m, int n, double2* C, int ldc,
                 hipStream_t s) {
    const int64_t N = c.N;
    const int tiles = ((m + kGM - 1) / kGM) * ((n + kGN - 1) / kGN);
    int chunks = std::max(1, std::min<int>(static_cast<int>((N + kGR - 1) / kGR), 1024 / tiles));
    int64_t rows = ((N + chunks - 1) / chunks + kGR - 1) / kGR * kGR;
    chunks = static_cast<int>((N + rows - 1) / rows);
    c.s_gram.alloc(std::max<size_t>(c.s_gram.n, static_cast<size_t>(chunks) * m * n));
    hipLaunchKernelGGL(gram_partial_kernel, dim3((m + kGM - 1) / kGM, (n + kGN - 1) / kGN, chunks), dim3(256), 0, s, A,
                       lda, B, ldb, N, m, n, rows, c.s_gram.p);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((m * n + 255) / 256), dim3(256), 0, s, c.s_gram.p, m, n, chunks, C, ldc);
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
