// Products Y = H X with the resident FP32 matrix on the FP16 matrix cores, each FP32
// operand split into two FP16 parts (hi = fp16(v), lo = fp16(v - hi), 22 significant
// bits together) and the three leading partial products accumulated in FP32:
//   v w ~ vh wh + vh wl + vl wh   (the dropped vl wl is below 2^-22 relative).
// The FP16 rate (2.5 PFLOP/s dense) makes the 12 FP16 products per complex
// multiply-add cheaper than the 3 FP32 ones, so the kernel is bound by streaming H
// from HBM (8 bytes per complex element), not by the matrix cores.
//
// Scaling: H is multiplied by sH = 2^(14 - e) with max |H| < 2^e (a rigorous bound from
// the diagonal, the interaction tables and the band-vector norms), each vector column
// by its own power of two the same way, so that every hi part is a normal FP16 number
// below 2^14 and the lo parts keep 11 more bits; the reduction divides the scales out.
//
// Layout: H is Hermitian, so row i of H is the conjugate of column i, which is
// contiguous in the column-major matrix. A wave owns 16 TW output rows (TW tiles of 16)
// and all 32 vectors; lane (i = l % 16, g = l / 16) of tile t reads 8 consecutive
// complex entries of column row0 + 16 t + i per 32-k step as four 16-byte loads, load q
// covering k = 8 q + 2 g + {0, 1}, so that each load instruction reads 64 contiguous
// bytes per column. The vectors are staged in the same k order as FP16 fragments
// Xs[step][n][g] = {c_hi, c_lo, d_hi, d_lo} (x = c + i d, 64 bytes), so that a lane's
// B operand for one 16-vector tile is one contiguous 64-byte read.
// With A = conj(H(k, i)) = a - i b: Re y = a c + b d, Im y = a d - b c (three FP32
// accumulators per tile: a c + b d, a d, b c).
// Columns are split across workgroups (split-K, XCD-aware as in the FP32 kernel); the
// last step of the last split may read up to 31 entries past the end of a column: the
// staged vectors are zero there and the matrix allocation carries 32 zero entries of
// padding past its end.
#include <cmath>

#include "context.hpp"
#include "cplx.hpp"
#include "mfma.hpp"

namespace xatu_dev {

namespace {

constexpr int kVec = kBlockVectors;
constexpr int kSplitTiles = 8;      // 16-row tiles per wave: 128 rows per wave, 512 per workgroup
constexpr int kMaxChunks = 64;      // row chunks of the column-maximum pass

struct XFrag16 {
    f16x8 ch, cl, dh, dl;
};
struct HFrag16 {
    f16x8 ah, al, bh, bl;
};

// Largest |value| bounds for the scale of H: per block, the maxima of |M(q)| over the
// direct table, of |M(Q)| over the exchange table and of the squared norms of the band
// vectors of both stacks.
__global__ void operator_bounds_kernel(const double2* __restrict__ ft, int64_t nft, const double2* __restrict__ ftq,
                                       int nftq, const double2* __restrict__ evk, const double2* __restrict__ evkq,
                                       int64_t ncols, int norb, double* __restrict__ partial) {
    __shared__ double red[3][256];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (int64_t i = t0; i < nft; i += stride) m0 = fmax(m0, hypot(ft[i].x, ft[i].y));
    for (int64_t i = t0; i < nftq; i += stride) m1 = fmax(m1, hypot(ftq[i].x, ftq[i].y));
    for (int64_t i = t0; i < 2 * ncols; i += stride) {
        const double2* v = (i < ncols ? evk : evkq) + (i % ncols) * norb;
        double s = 0.0;
        for (int a = 0; a < norb; ++a) s += v[a].x * v[a].x + v[a].y * v[a].y;
        m2 = fmax(m2, s);
    }
    red[0][threadIdx.x] = m0;
    red[1][threadIdx.x] = m1;
    red[2][threadIdx.x] = m2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int j = 0; j < 3; ++j) red[j][threadIdx.x] = fmax(red[j][threadIdx.x], red[j][threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// partial[n kMaxChunks + chunk] = max |Re|, |Im| of X[rows of chunk, n]; grid (chunks, 32).
__global__ void column_max_kernel(const double2* __restrict__ X, int64_t N, int nvec, int64_t rows,
                                  float* __restrict__ partial) {
    __shared__ double red[256];
    const int chunk = blockIdx.x, n = blockIdx.y;
    double m = 0.0;
    if (n < nvec) {
        const int64_t r1 = min(N, (chunk + 1) * rows);
        for (int64_t r = chunk * rows + threadIdx.x; r < r1; r += blockDim.x) {
            const double2 v = X[r + n * N];
            m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
        }
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    // rounded up so that the float maximum never understates the double one
    if (threadIdx.x == 0) partial[n * kMaxChunks + chunk] = __double2float_ru(red[0]);
}

// Power-of-two scale of each vector column (max |x| s < 2^14) and the inverse of the
// combined scale applied by the reduction.
__global__ void vector_scale_kernel(const float* __restrict__ partial, int chunks, float sH, float* __restrict__ s,
                                    double* __restrict__ inv) {
    const int n = threadIdx.x;
    if (n >= kVec) return;
    float m = 0.0f;
    for (int c = 0; c < chunks; ++c) m = fmaxf(m, partial[n * kMaxChunks + c]);
    int e = 0;
    if (m > 0.0f) frexpf(m, &e);
    const float sc = m > 0.0f ? ldexpf(1.0f, 14 - e) : 1.0f;
    s[n] = sc;
    inv[n] = 1.0 / (static_cast<double>(sc) * static_cast<double>(sH));
}

// Xs[(step 32 + n) 4 + g]: element e = 2 q + r of the fragment holds k = 32 step + 8 q + 2 g + r.
__global__ void stage_split_kernel(const double2* __restrict__ X, int64_t N, int nvec, int64_t nsteps,
                                   const float* __restrict__ s, XFrag16* __restrict__ Xs) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= nsteps * kVec * 4) return;
    const int g = static_cast<int>(idx & 3);
    const int n = static_cast<int>((idx >> 2) & (kVec - 1));
    const int64_t step = idx >> 7;
    const float sc = n < nvec ? s[n] : 0.0f;
    XFrag16 f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int64_t k = 32 * step + 8 * (e >> 1) + 2 * g + (e & 1);
        float c = 0.0f, d = 0.0f;
        if (n < nvec && k < N) {
            const double2 v = X[k + n * N];
            c = static_cast<float>(v.x * sc);
            d = static_cast<float>(v.y * sc);
        }
        const _Float16 ch = static_cast<_Float16>(c), dh = static_cast<_Float16>(d);
        f.ch[e] = ch;
        f.cl[e] = static_cast<_Float16>(c - static_cast<float>(ch));
        f.dh[e] = dh;
        f.dl[e] = static_cast<_Float16>(d - static_cast<float>(dh));
    }
    Xs[idx] = f;
}

__device__ inline HFrag16 split_h(const float4 (&r)[4], float sH) {
    HFrag16 f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float a0 = r[q].x * sH, b0 = r[q].y * sH, a1 = r[q].z * sH, b1 = r[q].w * sH;
        const _Float16 ha0 = static_cast<_Float16>(a0), hb0 = static_cast<_Float16>(b0);
        const _Float16 ha1 = static_cast<_Float16>(a1), hb1 = static_cast<_Float16>(b1);
        f.ah[2 * q] = ha0;
        f.ah[2 * q + 1] = ha1;
        f.bh[2 * q] = hb0;
        f.bh[2 * q + 1] = hb1;
        f.al[2 * q] = static_cast<_Float16>(a0 - static_cast<float>(ha0));
        f.al[2 * q + 1] = static_cast<_Float16>(a1 - static_cast<float>(ha1));
        f.bl[2 * q] = static_cast<_Float16>(b0 - static_cast<float>(hb0));
        f.bl[2 * q + 1] = static_cast<_Float16>(b1 - static_cast<float>(hb1));
    }
    return f;
}

// 1-D grid of row_blocks * splits workgroups, block 256; split = id % splits (splits a
// multiple of 8). Partial sums (scaled) to Yp[split][row][32].
template <int TW>
__global__ __launch_bounds__(256, 1) void split_product_kernel(const float2* __restrict__ H,
                                                               const XFrag16* __restrict__ Xs, int64_t N,
                                                               int64_t kchunk, int splits, float sH,
                                                               float2* __restrict__ Yp) {
    const int sp = blockIdx.x % splits;
    const int rb = blockIdx.x / splits;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int64_t row0 = static_cast<int64_t>(rb) * (64 * TW) + wave * (16 * TW);
    const int64_t kbeg = static_cast<int64_t>(sp) * kchunk;
    const int64_t kend = min(N, kbeg + kchunk);
    const int64_t nsteps = kend > kbeg ? (kend - kbeg + 31) / 32 : 0;
    f32x4 acc[TW][2][3];
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int a = 0; a < 3; ++a) acc[t][u][a] = f32x4{0, 0, 0, 0};
    // rows past N are clamped to row N-1 (their outputs are never stored)
    const float4* hcol[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        const int64_t i = row0 + 16 * t + li;
        hcol[t] = reinterpret_cast<const float4*>(H + (i < N ? i : N - 1) * N + kbeg + 2 * g);
    }
    const XFrag16* xb = Xs + ((kbeg / 32) * kVec + li) * 4 + g;
    for (int64_t st = 0; st < nsteps; ++st) {
        float4 raw[TW][4];
        XFrag16 xf[2];
#pragma unroll
        for (int t = 0; t < TW; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) raw[t][q] = hcol[t][st * 16 + 4 * q];
#pragma unroll
        for (int u = 0; u < 2; ++u) xf[u] = xb[(st * kVec + 16 * u) * 4];
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            const HFrag16 h = split_h(raw[t], sH);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const XFrag16& x = xf[u];
                f32x4 re = acc[t][u][0], ad = acc[t][u][1], bc = acc[t][u][2];
                re = mma_f16(h.al, x.ch, re);
                re = mma_f16(h.ah, x.cl, re);
                re = mma_f16(h.ah, x.ch, re);
                re = mma_f16(h.bl, x.dh, re);
                re = mma_f16(h.bh, x.dl, re);
                re = mma_f16(h.bh, x.dh, re);
                ad = mma_f16(h.al, x.dh, ad);
                ad = mma_f16(h.ah, x.dl, ad);
                ad = mma_f16(h.ah, x.dh, ad);
                bc = mma_f16(h.bl, x.ch, bc);
                bc = mma_f16(h.bh, x.cl, bc);
                bc = mma_f16(h.bh, x.ch, bc);
                acc[t][u][0] = re;
                acc[t][u][1] = ad;
                acc[t][u][2] = bc;
            }
        }
    }
    float2* out = Yp + static_cast<int64_t>(sp) * N * kVec;
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = row0 + 16 * t + 4 * g + v;
            if (r < N) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    out[r * kVec + 16 * u + li] =
                        make_float2(acc[t][u][0][v], acc[t][u][1][v] - acc[t][u][2][v]);
            }
        }
}

}  // namespace

void split_product_scale(Context& c) {
    const size_t nft = static_cast<size_t>(c.lat.natoms) * c.lat.natoms * c.lat.ntab;
    const int nftq = c.lat.natoms * c.lat.natoms;
    const int64_t ncols = static_cast<int64_t>(c.nk) * c.nb;
    constexpr int kBlocks = 64;
    DevBuf<double> part;
    part.alloc(3 * kBlocks);
    hipLaunchKernelGGL(operator_bounds_kernel, dim3(kBlocks), dim3(256), 0, c.stream, c.d_ft.p,
                       static_cast<int64_t>(nft), c.d_ftq.p, nftq, c.d_evec_k.p, c.d_evec_kq.p, ncols, c.lat.norb,
                       part.p);
    XATU_HIP(hipGetLastError());
    std::vector<double> h(3 * kBlocks);
    XATU_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    XATU_HIP(hipStreamSynchronize(c.stream));
    double mM = 0.0, mQ = 0.0, cn = 0.0, md = 0.0;
    for (int b = 0; b < kBlocks; ++b) {
        mM = std::max(mM, h[3 * b]);
        mQ = std::max(mQ, h[3 * b + 1]);
        cn = std::max(cn, h[3 * b + 2]);
    }
    for (double d : c.h_diag) md = std::max(md, std::fabs(d));
    if (!c.inter.exchange) mQ = 0.0;
    // |D_ij| <= max|M| (sum_a |rho^c_a|)(sum_b |rho^v_b|) <= max|M| cn^2, likewise |X_ij|;
    // 1e-6 relative margin for the FP32 rounding of the stored matrix
    const double bound = (md + (mM + mQ) * cn * cn) * (1.0 + 1e-6);
    int e = 0;
    if (bound > 0.0 && std::isfinite(bound)) std::frexp(bound, &e);
    c.h16_scale = (bound > 0.0 && std::isfinite(bound)) ? static_cast<float>(std::ldexp(1.0, 14 - e)) : 0.0f;
    if (!std::isnormal(c.h16_scale)) c.h16_scale = 0.0f;
}

bool split_product_available(const Context& c) {
    return c.built32 && !c.fp32_arith_only && c.h16_scale > 0.0f && (c.N % 2) == 0;
}

void launch_matvec_split16(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    constexpr int TW = kSplitTiles;
    const int64_t N = c.N;
    const int rows_wg = 64 * TW;
    const int row_blocks = static_cast<int>((N + rows_wg - 1) / rows_wg);
    // about four workgroups per CU (one wave per SIMD each); more splits only add partial sums
    int splits = std::max(1, std::min(64, (1024 + row_blocks - 1) / row_blocks));
    splits = (splits + kXcds - 1) / kXcds * kXcds;
    const int64_t kchunk = ((N + splits - 1) / splits + 31) / 32 * 32;
    const int64_t nsteps = (N + 31) / 32;
    const int64_t rows = (N + kMaxChunks - 1) / kMaxChunks;
    const int chunks = static_cast<int>((N + rows - 1) / rows);
    const size_t xs_bytes = static_cast<size_t>(nsteps) * kVec * 4 * sizeof(XFrag16);
    c.s_xs16.alloc(std::max(c.s_xs16.n, (xs_bytes + sizeof(float4) - 1) / sizeof(float4)));
    c.s_yp32.alloc(std::max<size_t>(c.s_yp32.n, static_cast<size_t>(splits) * N * kVec));
    c.s_colmax.alloc(static_cast<size_t>(kVec) * kMaxChunks);
    c.s_xscale.alloc(kVec);
    c.s_xinv.alloc(kVec);
    XFrag16* xs = reinterpret_cast<XFrag16*>(c.s_xs16.p);
    const float sH = c.h16_scale;
    for (int v0 = 0; v0 < nvec; v0 += kVec) {
        const int nv = std::min(kVec, nvec - v0);
        const double2* Xb = X + static_cast<int64_t>(v0) * N;
        hipLaunchKernelGGL(column_max_kernel, dim3(chunks, kVec), dim3(256), 0, s, Xb, N, nv, rows, c.s_colmax.p);
        hipLaunchKernelGGL(vector_scale_kernel, dim3(1), dim3(64), 0, s, c.s_colmax.p, chunks, sH, c.s_xscale.p,
                           c.s_xinv.p);
        hipLaunchKernelGGL(stage_split_kernel, dim3((nsteps * kVec * 4 + 255) / 256), dim3(256), 0, s, Xb, N, nv,
                           nsteps, c.s_xscale.p, xs);
        hipLaunchKernelGGL((split_product_kernel<TW>), dim3(row_blocks * splits), dim3(256), 0, s, c.d_H32.p, xs, N,
                           kchunk, splits, sH, c.s_yp32.p);
        launch_reduce_partials(c.s_yp32.p, N, nv, splits, Y + static_cast<int64_t>(v0) * N, s, c.s_xinv.p);
    }
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
