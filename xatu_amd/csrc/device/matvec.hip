// Dense products Y = H X with the resident BSE matrix on the matrix cores, and the
// split-K Gram kernel C = A^H B used by the Davidson projections.
//
// Product: H is column-major complex (FP32: float2, FP64: double2), N x N. The
// vectors are staged transposed, Xt[j][n] (32 vectors, zero padded). Each wave owns
// 32 rows (two 16-row tiles) and all 32 vectors (two 16-column tiles) and walks the
// columns of H four at a time: lane l loads the complex H(row = l%16 + 16 t, col =
// j + l/16), so the 16x16x4 MFMA sees A[i][k] = Re H(i, j+k) (or Im, or Re + Im) and
// B[k][n] built from X(j+k, n). FP64 uses four real MFMA products per tile pair
// (Re Re, Im (-Im), Re Im, Im Re); FP32 uses the three-product form. The loads of
// several 4-column steps are issued before their MFMAs to keep more bytes in flight.
// Column ranges are split across workgroups (split-K, XCD-aware) to fill all 256 CUs;
// partial results go to a workspace and a reduction kernel sums them in a fixed order
// and writes Y in the caller's column-major layout.
#include "context.hpp"
#include "cplx.hpp"
#include "mfma.hpp"

namespace xatu_dev {

namespace {

constexpr int kVec = kBlockVectors;

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

// Tile parameters of the product: TW 16-row tiles per wave (4 waves per workgroup, so
// 64 TW rows per workgroup) and U 4-column steps whose loads are issued together.
// FP32 uses the 3-real-product complex multiply (Gauss): with H = a + ib and x = c + id,
// k1 = (a + b) c, k2 = a (d - c), k3 = b (c + d), Re = k1 - k3, Im = k1 + k2.
template <typename R>
struct ProductTiling;
template <>
struct ProductTiling<float> {
    static constexpr int TW = 2, U = 4;
    static constexpr bool G3M = true;
};
template <>
struct ProductTiling<double> {
    static constexpr int TW = 2, U = 4;
    static constexpr bool G3M = false;
};

template <typename R, bool G3M, int TW>
__device__ inline void product_step(const cx_t<R> (&h)[TW], const cx_t<R> (&x)[2],
                                    typename Mfma<R>::acc_t (&acc)[TW][2][G3M ? 3 : 2]) {
    using M = Mfma<R>;
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const cx_t<R> hv = h[t], xv = x[u];
            if (G3M) {
                acc[t][u][0] = M::mma(hv.x + hv.y, xv.x, acc[t][u][0]);
                acc[t][u][1] = M::mma(hv.x, xv.y - xv.x, acc[t][u][1]);
                acc[t][u][2] = M::mma(hv.y, xv.x + xv.y, acc[t][u][2]);
            } else {
                acc[t][u][0] = M::mma(hv.x, xv.x, acc[t][u][0]);
                acc[t][u][0] = M::mma(hv.y, -xv.y, acc[t][u][0]);
                acc[t][u][1] = M::mma(hv.x, xv.y, acc[t][u][1]);
                acc[t][u][1] = M::mma(hv.y, xv.x, acc[t][u][1]);
            }
        }
}

// 1-D grid of row_blocks * splits workgroups, block 256. Workgroup id -> (split = id %
// splits, row block = id / splits): with splits a multiple of 8 every split, and with
// it its slice of Xt, lives on one XCD and its L2. Partial sums to Yp[split][row][32].
// The main loop has no conditional loads: rows past N are clamped to row N-1 once (they
// only feed outputs that are never stored) and whole 4U-column groups are loaded
// unconditionally, so the compiler can issue all loads of a group together; only the
// column tail of a split is masked.
template <typename R>
__global__ __launch_bounds__(256, 3) void dense_product_kernel(const cx_t<R>* __restrict__ H, const cx_t<R>* __restrict__ Xt,
                                                            int64_t N, int64_t kchunk, int splits,
                                                            cx_t<R>* __restrict__ Yp) {
    using M = Mfma<R>;
    using acc_t = typename M::acc_t;
    using T = ProductTiling<R>;
    constexpr int TW = T::TW, U = T::U;
    constexpr bool G3 = T::G3M;
    constexpr int NACC = G3 ? 3 : 2;
    const int sp = blockIdx.x % splits;
    const int rb = blockIdx.x / splits;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t row0 = static_cast<int64_t>(rb) * (64 * TW) + wave * (16 * TW);
    const int64_t kbeg = static_cast<int64_t>(sp) * kchunk;
    const int64_t kend = min(N, kbeg + kchunk);
    const int li = lane & 15;  // row within a 16-row tile / vector within a 16-vector tile
    const int lk = lane >> 4;  // column offset 0..3 within a 4-column step

    acc_t acc[TW][2][NACC];
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int a = 0; a < NACC; ++a) acc[t][u][a] = acc_t{0, 0, 0, 0};

    const cx_t<R>* hrow[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        const int64_t r = row0 + 16 * t + li;
        hrow[t] = H + (r < N ? r : N - 1) + static_cast<int64_t>(lk) * N;
    }
    const cx_t<R>* xl = Xt + static_cast<int64_t>(lk) * kVec + li;

    const int64_t kmain = kbeg + (kend > kbeg ? (kend - kbeg) / (4 * U) * (4 * U) : 0);
    int64_t j = kbeg;
    for (; j < kmain; j += 4 * U) {
        cx_t<R> h[U][TW], x[U][2];
#pragma unroll
        for (int s = 0; s < U; ++s) {
            const int64_t col = j + 4 * s;
#pragma unroll
            for (int t = 0; t < TW; ++t) h[s][t] = hrow[t][col * N];
#pragma unroll
            for (int u = 0; u < 2; ++u) x[s][u] = xl[col * kVec + 16 * u];
        }
#pragma unroll
        for (int s = 0; s < U; ++s) product_step<R, G3, TW>(h[s], x[s], acc);
    }
    for (; j < kend; j += 4) {  // masked tail, fewer than 4U columns
        const bool cok = j + lk < kend;
        const int64_t col = cok ? j : kend - 1 - lk;  // keeps the lane's address inside the split
        cx_t<R> h[TW], x[2];
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            const cx_t<R> v = hrow[t][col * N];
            h[t] = cok ? v : mk(R(0), R(0));
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const cx_t<R> v = xl[col * kVec + 16 * u];
            x[u] = cok ? v : mk(R(0), R(0));
        }
        product_step<R, G3, TW>(h, x, acc);
    }
    cx_t<R>* out = Yp + static_cast<int64_t>(sp) * N * kVec;
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = row0 + 16 * t + M::out_row(lk, v);
            if (r < N) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    R re, im;
                    if (G3) {
                        re = acc[t][u][0][v] - acc[t][u][2][v];
                        im = acc[t][u][0][v] + acc[t][u][1][v];
                    } else {
                        re = acc[t][u][0][v];
                        im = acc[t][u][1][v];
                    }
                    out[r * kVec + 16 * u + li] = mk(re, im);
                }
            }
        }
}

// Y[r + n N] = sum_s Yp[s][r][n] for n < nvec, summed in the output precision in a fixed
// order. A workgroup owns 64 rows: the partial rows are read
// with consecutive threads on consecutive (row, vector) pairs, the sums are transposed
// through LDS and written column by column.
template <typename R, typename Rout>
__global__ __launch_bounds__(256) void reduce_product_kernel(const cx_t<R>* __restrict__ Yp, int64_t N, int nvec,
                                                             int splits, cx_t<Rout>* __restrict__ Y) {
    constexpr int kRows = 64;
    __shared__ cx_t<Rout> tile[kRows][kVec + 1];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kRows;
#pragma unroll
    for (int j = 0; j < kRows * kVec / 256; ++j) {
        const int f = threadIdx.x + 256 * j;
        const int rr = f / kVec, n = f % kVec;
        const int64_t r = r0 + rr;
        Rout sr = 0, si = 0;
        if (r < N && n < nvec) {
            for (int s = 0; s < splits; ++s) {
                const cx_t<R> v = Yp[(static_cast<int64_t>(s) * N + r) * kVec + n];
                sr += static_cast<Rout>(v.x);
                si += static_cast<Rout>(v.y);
            }
        }
        tile[rr][n] = mk(sr, si);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRows * kVec / 256; ++j) {
        const int f = threadIdx.x + 256 * j;
        const int rr = f % kRows, n = f / kRows;
        const int64_t r = r0 + rr;
        if (r < N && n < nvec) Y[r + static_cast<int64_t>(n) * N] = tile[rr][n];
    }
}

// C (N x nb) = alpha A B + beta C (beta = 0: C is not read) for a tall A (N x m,
// column-major, leading dimension lda) and a small B (m x nb, ldb), nb <= 32, FP64
// complex: 64 rows per workgroup, one wave per kBgCols columns. The B entries a wave needs are
// the same for all its lanes, so they are read with uniform addresses (scalar loads) and
// feed the FMAs as scalar operands: no LDS traffic and no barriers. Four columns per wave
// (eight waves per workgroup) measured faster than two, eight or sixteen: the kernel is
// bound by loads in flight, and more waves keep more of them in flight.
constexpr int kBgCols = 4, kBgWaves = kVec / kBgCols;
static_assert(kVec % kBgCols == 0, "whole waves per column block");
__global__ __launch_bounds__(64 * kBgWaves) void block_gemm_kernel(const double2* __restrict__ A, int lda,
                                                         const double2* __restrict__ B, int ldb, int64_t N, int m,
                                                         int nb, double alpha, double beta, double2* __restrict__ C,
                                                         int ldc) {
    const int r = threadIdx.x & 63;
    const int cg = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 64 + r;
    const int64_t ia = i < N ? i : N - 1;  // rows past N load row N-1 and are not stored
    const int jn = nb - kBgCols * cg;      // columns of this wave present (uniform)
    double2 acc[kBgCols];
#pragma unroll
    for (int jj = 0; jj < kBgCols; ++jj) acc[jj] = make_double2(0.0, 0.0);
    if (jn > 0) {
        const double2* Bc = B + static_cast<int64_t>(kBgCols * cg) * ldb;
        const double2* Ar = A + ia;
        // unrolled so that several column loads of A are in flight together
#pragma unroll 4
        for (int k = 0; k < m; ++k) {
            const double2 a = Ar[static_cast<int64_t>(k) * lda];
#pragma unroll
            for (int jj = 0; jj < kBgCols; ++jj) {
                // a column past nb repeats the last one present; its sum is not stored
                const double2 b = Bc[k + static_cast<int64_t>(min(jj, jn - 1)) * ldb];
                cfma(acc[jj], a, b);
            }
        }
    }
    if (i < N) {
#pragma unroll
        for (int jj = 0; jj < kBgCols; ++jj) {
            const int j = kBgCols * cg + jj;
            if (j < nb) {
                double2* cp = C + i + static_cast<int64_t>(j) * ldc;
                double2 v = make_double2(alpha * acc[jj].x, alpha * acc[jj].y);
                if (beta != 0.0) {
                    const double2 o = *cp;
                    v.x += beta * o.x;
                    v.y += beta * o.y;
                }
                *cp = v;
            }
        }
    }
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
    // the next 16-row slab is loaded into registers while the current one is used
    static_assert(kGR * kGM == 4 * 256 && kGR * kGN == 2 * 256, "one slab = 4 A and 2 B entries per thread");
    double2 ra[4], rb[2];
    auto fetch = [&](int64_t r0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = threadIdx.x + 256 * q, rr = e % kGR, cc = e / kGR;
            const int64_t r = r0 + rr;
            ra[q] = (r < rend && m0 + cc < m) ? A[r + static_cast<int64_t>(m0 + cc) * lda] : mk(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + 256 * q, rr = e % kGR, cc = e / kGR;
            const int64_t r = r0 + rr;
            rb[q] = (r < rend && n0 + cc < n) ? B[r + static_cast<int64_t>(n0 + cc) * ldb] : mk(0.0, 0.0);
        }
    };
    if (rbeg < rend) fetch(rbeg);
    for (int64_t r0 = rbeg; r0 < rend; r0 += kGR) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = threadIdx.x + 256 * q;
            sA[e % kGR][e / kGR] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + 256 * q;
            sB[e % kGR][e / kGR] = rb[q];
        }
        __syncthreads();
        if (r0 + kGR < rend) fetch(r0 + kGR);
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

// C[i + j ldc] = sum over chunks of P[chunk][i + j m]. One workgroup of 1024 threads
// per 64 outputs: 16 slices each sum every 16th chunk, then the slices are added in
// a fixed order, so the result does not depend on scheduling.
constexpr int kRedOut = 64, kRedSlices = 16;
__global__ __launch_bounds__(kRedOut* kRedSlices) void gram_reduce_kernel(const double2* __restrict__ P, int m, int n,
                                                                          int chunks, double2* __restrict__ C,
                                                                          int ldc) {
    __shared__ double2 part[kRedSlices][kRedOut];
    const int o = threadIdx.x % kRedOut, sl = threadIdx.x / kRedOut;
    const int idx = blockIdx.x * kRedOut + o;
    const int64_t mn = static_cast<int64_t>(m) * n;
    double sr = 0.0, si = 0.0;
    if (idx < mn) {
#pragma unroll 4
        for (int c = sl; c < chunks; c += kRedSlices) {
            const double2 v = P[static_cast<int64_t>(c) * mn + idx];
            sr += v.x;
            si += v.y;
        }
    }
    part[sl][o] = mk(sr, si);
    __syncthreads();
    if (sl == 0 && idx < mn) {
        for (int k = 1; k < kRedSlices; ++k) {
            sr += part[k][o].x;
            si += part[k][o].y;
        }
        const int i = idx % m, j = idx / m;
        C[i + static_cast<int64_t>(j) * ldc] = mk(sr, si);
    }
}

// Column splits of the dense product: enough workgroups to fill 256 CUs a few times over,
// a multiple of the XCD count, at most max_splits (>= 1).
template <typename R>
int product_splits(int64_t N, int max_splits) {
    using T = ProductTiling<R>;
    const int rows_wg = 64 * T::TW;
    const int row_blocks = static_cast<int>((N + rows_wg - 1) / rows_wg);
    int splits = std::max(1, std::min(64, (2048 + row_blocks - 1) / row_blocks));
    splits = (splits + kXcds - 1) / kXcds * kXcds;
    return std::max(1, std::min(splits, max_splits));
}

// Y = H X through the staging buffer xt (N x 32) and the partial sums yp (splits x N x 32).
template <typename Rin, typename R, typename Rout>
void dense_product_into(int64_t N, const cx_t<R>* H, const cx_t<Rin>* X, cx_t<Rout>* Y, int nvec, hipStream_t s,
                        cx_t<R>* xt, cx_t<R>* yp, int splits) {
    using T = ProductTiling<R>;
    const int rows_wg = 64 * T::TW;
    const int row_blocks = static_cast<int>((N + rows_wg - 1) / rows_wg);
    const int step = 4 * T::U;
    const int64_t kchunk = ((N + splits - 1) / splits + step - 1) / step * step;
    // every split owns a (possibly empty) column range and writes its partial rows
    for (int v0 = 0; v0 < nvec; v0 += kVec) {
        const int nv = std::min(kVec, nvec - v0);
        const int64_t tot = N * kVec;
        hipLaunchKernelGGL((stage_vectors_kernel<Rin, R>), dim3((tot + 255) / 256), dim3(256), 0, s, X + v0 * N, N, nv,
                           xt);
        hipLaunchKernelGGL((dense_product_kernel<R>), dim3(row_blocks * splits), dim3(256), 0, s, H, xt, N, kchunk,
                           splits, yp);
        hipLaunchKernelGGL((reduce_product_kernel<R, Rout>), dim3((N + 63) / 64), dim3(256), 0, s, yp, N, nv, splits,
                           Y + v0 * N);
    }
    XATU_HIP(hipGetLastError());
}

template <typename Rin, typename R, typename Rout>
void dense_product(const Context& c, const cx_t<R>* H, const cx_t<Rin>* X, cx_t<Rout>* Y, int nvec, hipStream_t s,
                   DevBuf<cx_t<R>>& xt, DevBuf<cx_t<R>>& yp) {
    const int64_t N = c.N;
    const int splits = product_splits<R>(N, 64);
    xt.alloc(std::max<size_t>(xt.n, static_cast<size_t>(N) * kVec));
    yp.alloc(std::max<size_t>(yp.n, static_cast<size_t>(splits) * N * kVec));
    dense_product_into<Rin, R, Rout>(N, H, X, Y, nvec, s, xt.p, yp.p, splits);
}

}  // namespace

void launch_matvec64(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    dense_product<double, double, double>(c, c.d_H64.p, X, Y, nvec, s, c.s_xt64, c.s_yp64);
}

int64_t matvec64_scratch_elems(int64_t N, int max_splits) {
    return N * kVec * (1 + product_splits<double>(N, max_splits));
}

void launch_matvec64_scratch(const Context& c, const double2* X, double2* Y, int nvec, double2* scratch,
                             int64_t scratch_elems, hipStream_t s) {
    const int64_t N = c.N;
    // as many splits as the scratch holds beside the staged vectors (same kernels and, when
    // they fit, the same splits as the allocating product)
    const int max_splits = static_cast<int>(std::min<int64_t>(64, scratch_elems / (N * kVec) - 1));
    if (max_splits < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "scratch too small for the product");
    const int splits = product_splits<double>(N, max_splits);
    dense_product_into<double, double, double>(N, c.d_H64.p, X, Y, nvec, s, scratch, scratch + N * kVec, splits);
}

// Product with the FP32 matrix: on the FP16 matrix cores with split operands where the
// matrix layout allows it (even N), else staged in FP32 and accumulated in FP32 on the
// FP32 matrix cores; widened to FP64 on output either way.
void launch_matvec32(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    if (split_product_available(c)) {
        launch_matvec_split16(c, X, Y, nvec, s);
        return;
    }
    complete_matrix32(c);
    dense_product<double, float, double>(c, c.d_H32.p, X, Y, nvec, s, c.s_xt32, c.s_yp32);
}

void launch_stage_vectors(const double2* X, int64_t N, int nvec, double2* Xt, hipStream_t s) {
    hipLaunchKernelGGL((stage_vectors_kernel<double, double>), dim3((N * kVec + 255) / 256), dim3(256), 0, s, X, N,
                       nvec, Xt);
}
void launch_stage_vectors(const double2* X, int64_t N, int nvec, float2* Xt, hipStream_t s) {
    hipLaunchKernelGGL((stage_vectors_kernel<double, float>), dim3((N * kVec + 255) / 256), dim3(256), 0, s, X, N,
                       nvec, Xt);
}
void launch_reduce_partials(const double2* Yp, int64_t N, int nvec, int splits, double2* Y, hipStream_t s) {
    hipLaunchKernelGGL((reduce_product_kernel<double, double>), dim3((N + 63) / 64), dim3(256), 0, s, Yp, N, nvec,
                       splits, Y);
}
void launch_reduce_partials(const float2* Yp, int64_t N, int nvec, int splits, double2* Y, hipStream_t s) {
    hipLaunchKernelGGL((reduce_product_kernel<float, double>), dim3((N + 63) / 64), dim3(256), 0, s, Yp, N, nvec,
                       splits, Y);
}

void launch_block_gemm(const double2* A, int lda, const double2* B, int ldb, int64_t N, int m, int nb, double alpha,
                       double beta, double2* C, int ldc, hipStream_t s) {
    for (int j0 = 0; j0 < nb; j0 += kVec) {
        const int nj = std::min(kVec, nb - j0);
        hipLaunchKernelGGL(block_gemm_kernel, dim3(static_cast<unsigned>((N + 63) / 64)), dim3(64 * kBgWaves), 0, s, A, lda,
                           B + static_cast<int64_t>(j0) * ldb, ldb, N, m, nj, alpha, beta,
                           C + static_cast<int64_t>(j0) * ldc, ldc);
    }
    XATU_HIP(hipGetLastError());
}

void launch_gram(Context& c, const double2* A, int lda, const double2* B, int ldb, int m, int n, double2* C, int ldc,
                 hipStream_t s) {
    const int64_t N = c.N;
    const int tiles = ((m + kGM - 1) / kGM) * ((n + kGN - 1) / kGN);
    int chunks = std::max(1, std::min<int>(static_cast<int>((N + kGR - 1) / kGR), 512 / tiles));
    int64_t rows = ((N + chunks - 1) / chunks + kGR - 1) / kGR * kGR;
    chunks = static_cast<int>((N + rows - 1) / rows);
    c.s_gram.alloc(std::max<size_t>(c.s_gram.n, static_cast<size_t>(chunks) * m * n));
    hipLaunchKernelGGL(gram_partial_kernel, dim3((m + kGM - 1) / kGM, (n + kGN - 1) / kGN, chunks), dim3(256), 0, s, A,
                       lda, B, ldb, N, m, n, rows, c.s_gram.p);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((m * n + kRedOut - 1) / kRedOut), dim3(kRedOut * kRedSlices), 0, s,
                       c.s_gram.p, m, n, chunks, C, ldc);
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
