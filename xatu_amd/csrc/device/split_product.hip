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
// Layout: H is Hermitian and column-major, so row i of H is the conjugate of column i,
// which is contiguous. The product reads each distinct entry once: row block b (256 rows,
// one workgroup of 4 waves, 64 rows per wave) reads columns i of b from row k = 256 b down
// (its diagonal block and the blocks below it). From one load it forms
//   direct blocks      y_i += sum_k conj(H(k, i)) x_k   (rows of b, accumulated in registers)
//   transposed blocks  y_k += sum_i H(k, i) x_i         (k below the diagonal block, per step)
// so that the block-upper part of H is never read (a block-lower build never writes it).
// Lane (i = l % 16, g = l / 16) of tile t reads 8 consecutive complex entries of column
// row0 + 16 t + i per 32-k step as four 16-byte loads, load q covering k = 8 q + 2 g + {0, 1};
// the vectors are staged in the same k order as FP16 fragments Xs[step][n][g] =
// {c_hi, c_lo, d_hi, d_lo} (x = c + i d, 64 bytes). For the transposed blocks the split
// parts pass through a per-wave LDS buffer so that a lane holds one k and eight i, the
// vectors being staged a second time in that i order (Xt). Complex products use two FP32
// accumulators per tile, the negated imaginary parts of H supplying the minus signs.
// The columns are cut into chunks of 2048 (work items (b, chunk) for the chunks at or
// below b's diagonal block); the direct sums go to one slot per chunk, the transposed
// sums (the four waves' added in LDS) to one slot per row block, and a reduction adds
// the slots of each row in a fixed order. The last step of a column may read up to 31
// entries past its end: the staged vectors are zero there and the entries read are finite
// (32 zero entries of padding after the matrix, zeroed heads of the unbuilt columns).
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.hpp"
#include "cplx.hpp"
#include "mfma.hpp"

namespace xatu_dev {

namespace {

constexpr int kVec = kBlockVectors;
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

__device__ inline void split_entry(const double2* __restrict__ X, int64_t N, int nvec, int n, float sc, int64_t k,
                                   int e, XFrag16& f) {
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

// Xs[(step 32 + n) 4 + g]: element e = 2 q + r of the fragment holds k = 32 step + 8 q + 2 g + r.
// Xt[((G 2 + jj) 32 + n) 4 + g]: element e holds row 64 G + 32 jj + 16 (e & 1) + 4 g + (e >> 1).
// Rows past N are zero.
__global__ void stage_split_kernel(const double2* __restrict__ X, int64_t N, int nvec, int64_t nsteps,
                                   int64_t ngroups, const float* __restrict__ s, XFrag16* __restrict__ Xs,
                                   XFrag16* __restrict__ Xt) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t ns = nsteps * kVec * 4;
    if (idx >= ns + ngroups * 2 * kVec * 4) return;
    const int g = static_cast<int>(idx & 3);
    const int n = static_cast<int>((idx >> 2) & (kVec - 1));
    const float sc = n < nvec ? s[n] : 0.0f;
    XFrag16 f;
    if (idx < ns) {
        const int64_t step = idx >> 7;
#pragma unroll
        for (int e = 0; e < 8; ++e) split_entry(X, N, nvec, n, sc, 32 * step + 8 * (e >> 1) + 2 * g + (e & 1), e, f);
        Xs[idx] = f;
    } else {
        const int64_t it = idx - ns;
        const int jj = static_cast<int>((it >> 7) & 1);
        const int64_t G = it >> 8;
#pragma unroll
        for (int e = 0; e < 8; ++e)
            split_entry(X, N, nvec, n, sc, 64 * G + 32 * jj + 16 * (e & 1) + 4 * g + (e >> 1), e, f);
        Xt[it] = f;
    }
}

// The four loads of a tile step already hold the parts (a hi, a lo, b hi, b lo) when the
// matrix is stored split (the FP32 build writes that layout for this product).
__device__ inline HFrag16 presplit_h(const float4 (&r)[4]) {
    HFrag16 f;
    f.ah = __builtin_bit_cast(f16x8, r[0]);
    f.al = __builtin_bit_cast(f16x8, r[1]);
    f.bh = __builtin_bit_cast(f16x8, r[2]);
    f.bl = __builtin_bit_cast(f16x8, r[3]);
    return f;
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

// Work items of the product: (row block b, column chunk c) pairs with c >= the chunk
// holding the diagonal block of b, so that only the lower block triangle of H (and the
// diagonal blocks) is read. One workgroup of 4 waves owns the kHermRows rows of a row
// block, 16 kHermTiles rows per wave.
constexpr int kHermTiles = 4;
constexpr int kHermRows = 4 * 16 * kHermTiles;
static_assert(kHermRows == kSplitRowBlock, "row block of the split product");
constexpr int kHermChunk = 4096;   // columns per work item, a multiple of kHermRows
constexpr int kTrRow = 72;         // FP16 entries per row of the transposition buffer (64 + 8 padding)
constexpr int kCbRow = 34;         // complex entries per row of the combine buffer (32 + 2 padding)

struct HermLds {
    _Float16 tr[4][4][32 * kTrRow];  // per wave, per part (a hi, a lo, b hi, b lo): [rho][t / 2][li][t & 1]
    float2 cb[2][4][32][kCbRow];     // per slot (step parity), per wave: transposed-block partial sums
};

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// Orders the LDS accesses of one wave (written by some lanes, read by others) without
// waiting for its global loads.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One work item (b, c), grid = number of items, block 256.
//  Direct blocks: y_i += sum_k conj(H(k, i)) x_k for the rows i of b and the columns k of
//  the item (k >= first row of b), read down the columns of H as in the full product;
//  partial sums to Pn[c][i][32].
//  Transposed blocks: for k past the diagonal block, y_k += sum_{i in b} H(k, i) x_i with the
//  same loaded entries. Each wave's split FP16 parts go through its LDS buffer so that a
//  lane holds one k and eight i (rho = 8 q + 4 r + g orders the k of a step); the four waves'
//  sums over their 64 rows are combined in LDS and stored per step to Pt[b][k][32].
// Xt holds the vectors in the i order of the transposed fragments: element e of lane
// (n, g) of fragment jj of the wave's 64 rows holds row 32 jj + 16 (e & 1) + 4 g + (e >> 1).
template <int TW, bool PRESPLIT>
__global__ __launch_bounds__(256, 1) void herm_product_kernel(const float2* __restrict__ H,
                                                              const XFrag16* __restrict__ Xs,
                                                              const XFrag16* __restrict__ Xt,
                                                              const int2* __restrict__ items, int64_t N, float sH,
                                                              float2* __restrict__ Pn, float2* __restrict__ Pt) {
    static_assert(TW == 4, "the transposed fragments assume 64 rows per wave");
    __shared__ HermLds lds;
    const int2 bc = items[blockIdx.x];
    const int b = bc.x, c = bc.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int64_t rb0 = static_cast<int64_t>(b) * kHermRows;
    const int64_t row0 = rb0 + wave * (16 * TW);
    const int64_t cbeg = static_cast<int64_t>(c) * kHermChunk;
    const int64_t kend = min(N, cbeg + kHermChunk);
    const bool diag = cbeg <= rb0;
    const int64_t kbeg = diag ? rb0 : cbeg;
    const int nd = diag ? static_cast<int>((min(kend, rb0 + kHermRows) - rb0 + 31) / 32) : 0;
    const int64_t klow = diag ? rb0 + kHermRows : cbeg;
    const int nl = kend > klow ? static_cast<int>((kend - klow + 31) / 32) : 0;
    const int nsteps = nd + nl;

    f32x4 acc[TW][2][2];
    f32x4 acct[2][2][2];
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u][0] = acc[t][u][1] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < 2; ++u) acct[h][u][0] = acct[h][u][1] = f32x4{0, 0, 0, 0};
    // rows past N are clamped to row N-1 (their outputs are never stored, their vector
    // entries are staged as zero)
    const float4* hcol[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) {
        const int64_t i = row0 + 16 * t + li;
        hcol[t] = reinterpret_cast<const float4*>(H + (i < N ? i : N - 1) * N) + g;
    }
    XFrag16 xi[2][2];
    {
        const int64_t G = row0 / 64;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int u = 0; u < 2; ++u) xi[jj][u] = Xt[((G * 2 + jj) * kVec + 16 * u + li) * 4 + g];
    }
    _Float16* trw = &lds.tr[wave][0][0];
    // two steps of H in flight: step st + 2 is loaded into the registers of step st as soon
    // as these are split
    auto load = [&](float4 (&rb)[TW][4], XFrag16 (&xb)[2], int st) {
        const int64_t k0 = st < nd ? kbeg + 32 * st : klow + 32 * (st - nd);
#pragma unroll
        for (int t = 0; t < TW; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) rb[t][q] = hcol[t][k0 / 2 + 4 * q];
#pragma unroll
        for (int u = 0; u < 2; ++u) xb[u] = Xs[((k0 / 32) * kVec + 16 * u + li) * 4 + g];
    };
    // sums of one transposed step over the four waves, in wave order; wave w stores rows
    // 8 w .. 8 w + 7 of the step
    auto combine = [&](int slot, int64_t kc) {
        float2* out = Pt + static_cast<int64_t>(b) * N * kVec;
        float2 part[4][4];
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int f = lane + 64 * it;
            const int row = 8 * wave + (f >> 5), n = f & 31;
#pragma unroll
            for (int w = 0; w < 4; ++w) part[it][w] = lds.cb[slot][w][row][n];
        }
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int f = lane + 64 * it;
            const int row = 8 * wave + (f >> 5), n = f & 31;
            float2 sum = part[it][0];
#pragma unroll
            for (int w = 1; w < 4; ++w) {
                sum.x += part[it][w].x;
                sum.y += part[it][w].y;
            }
            const int64_t k = kc + row;
            if (k < N) out[k * kVec + n] = sum;
        }
    };
    // body(st) with st == nsteps (odd step counts) only reloads; loads past the last step
    // repeat the last step, so that every register set is reloaded unconditionally
    auto body = [&](float4 (&rb)[TW][4], XFrag16 (&xb)[2], int st) {
        const bool valid = st < nsteps;
        const bool low = valid && st >= nd;
        const int64_t k0 = st < nd ? kbeg + 32 * st : klow + 32 * (st - nd);
        const int sn = min(st + 2, nsteps - 1);
        const int64_t kn = sn < nd ? kbeg + 32 * sn : klow + 32 * (sn - nd);
        if (low) wave_lds_sync();
        // tiles in pairs: split, reload their registers with step st + 2, store the parts to
        // the transposition buffer, direct-block products
#pragma unroll
        for (int tp = 0; tp < TW / 2; ++tp) {
            HFrag16 hf[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) hf[j] = PRESPLIT ? presplit_h(rb[2 * tp + j]) : split_h(rb[2 * tp + j], sH);
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) rb[2 * tp + j][q] = hcol[2 * tp + j][kn / 2 + 4 * q];
            if (low) {
                // element e = 2 q + r of a fragment is k = 8 q + 2 g + r, row rho = 8 q + 4 r + g;
                // a row holds [tile pair][li][tile & 1]
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int rho = 8 * (e >> 1) + 4 * (e & 1) + g;
                    _Float16* dst = trw + rho * kTrRow + 32 * tp + 2 * li;
                    *reinterpret_cast<f16x2*>(dst) = f16x2{hf[0].ah[e], hf[1].ah[e]};
                    *reinterpret_cast<f16x2*>(dst + 32 * kTrRow) = f16x2{hf[0].al[e], hf[1].al[e]};
                    *reinterpret_cast<f16x2*>(dst + 64 * kTrRow) = f16x2{hf[0].bh[e], hf[1].bh[e]};
                    *reinterpret_cast<f16x2*>(dst + 96 * kTrRow) = f16x2{hf[0].bl[e], hf[1].bl[e]};
                }
            }
            // direct blocks: A = conj(H(k, i)) = a - i b, x = c + i d:
            // Re y += a c + b d, Im y += a d + (-b) c
#pragma unroll
            for (int j = 0; j < 2 && valid; ++j) {
                const int t = 2 * tp + j;
                const HFrag16& h = hf[j];
                const f16x8 nbh = -h.bh, nbl = -h.bl;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const XFrag16& x = xb[u];
                    f32x4 re = acc[t][u][0], im = acc[t][u][1];
                    re = mma_f16(h.al, x.ch, re);
                    re = mma_f16(h.ah, x.cl, re);
                    re = mma_f16(h.ah, x.ch, re);
                    re = mma_f16(h.bl, x.dh, re);
                    re = mma_f16(h.bh, x.dl, re);
                    re = mma_f16(h.bh, x.dh, re);
                    im = mma_f16(h.al, x.dh, im);
                    im = mma_f16(h.ah, x.dl, im);
                    im = mma_f16(h.ah, x.dh, im);
                    im = mma_f16(nbl, x.ch, im);
                    im = mma_f16(nbh, x.cl, im);
                    im = mma_f16(nbh, x.ch, im);
                    acc[t][u][0] = re;
                    acc[t][u][1] = im;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) xb[u] = Xs[((kn / 32) * kVec + 16 * u + li) * 4 + g];
        if (low) {
            wave_lds_sync();
            // transposed blocks: lane (m, g) holds row k(rho = 16 h + m) and the eight i of
            // fragment jj (tiles 2 jj, 2 jj + 1, li = 4 g .. 4 g + 3); A = H(k, i) = a + i b:
            // Re y += a c + (-b) d, Im y += a d + b c
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) {
                    const _Float16* src = trw + (16 * h + li) * kTrRow + 32 * jj + 8 * g;
                    const f16x8 ah = *reinterpret_cast<const f16x8*>(src);
                    const f16x8 al = *reinterpret_cast<const f16x8*>(src + 32 * kTrRow);
                    const f16x8 bh = *reinterpret_cast<const f16x8*>(src + 64 * kTrRow);
                    const f16x8 bl = *reinterpret_cast<const f16x8*>(src + 96 * kTrRow);
                    const f16x8 nbh = -bh, nbl = -bl;
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const XFrag16& x = xi[jj][u];
                        f32x4 re = acct[h][u][0], im = acct[h][u][1];
                        re = mma_f16(al, x.ch, re);
                        re = mma_f16(ah, x.cl, re);
                        re = mma_f16(ah, x.ch, re);
                        re = mma_f16(nbl, x.dh, re);
                        re = mma_f16(nbh, x.dl, re);
                        re = mma_f16(nbh, x.dh, re);
                        im = mma_f16(al, x.dh, im);
                        im = mma_f16(ah, x.dl, im);
                        im = mma_f16(ah, x.dh, im);
                        im = mma_f16(bl, x.ch, im);
                        im = mma_f16(bh, x.cl, im);
                        im = mma_f16(bh, x.ch, im);
                        acct[h][u][0] = re;
                        acct[h][u][1] = im;
                    }
                }
            // output row m = 4 g + v of fragment h is rho = 16 h + m, i.e. k = k0 + kl below
            const int slot = (st - nd) & 1;
            float2* cbw = &lds.cb[slot][wave][0][0];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int kl = 16 * h + 8 * (g >> 1) + (g & 1) + 2 * v;
                        cbw[kl * kCbRow + 16 * u + li] = make_float2(acct[h][u][0][v], acct[h][u][1][v]);
                        acct[h][u][0][v] = 0.0f;
                        acct[h][u][1][v] = 0.0f;
                    }
            __syncthreads();
            combine(slot, k0);
        }
    };
    float4 bufa[TW][4], bufb[TW][4];
    XFrag16 xa[2], xbb[2];
    load(bufa, xa, 0);
    load(bufb, xbb, min(1, nsteps - 1));
    for (int st = 0; st < nsteps; st += 2) {
        body(bufa, xa, st);
        body(bufb, xbb, st + 1);
    }

    float2* out = Pn + static_cast<int64_t>(c) * N * kVec;
#pragma unroll
    for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t r = row0 + 16 * t + 4 * g + v;
            if (r < N) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    out[r * kVec + 16 * u + li] = make_float2(acc[t][u][0][v], acc[t][u][1][v]);
            }
        }
}

// The same product for a matrix stored in the split layout, organised per wave: wave w of
// the workgroup takes the steps w, w + 4, ... of the item and, for each, all 256 columns i of
// the row block (eight groups of 32 = two tiles of 16). Direct sums accumulate per wave over
// its steps for all 256 rows (combined across the four waves once, at the end of the item);
// the transposed sums of a step are complete inside the wave (its 256 i) and are stored to
// Pt[b][k][32] at once, so the waves never wait for each other inside the loop. The vectors
// of the row block in the transposed-fragment order are staged in LDS once per item; each
// wave transposes its split parts through its own LDS buffer, one group at a time. Two
// register sets of two tiles alternate, so one group is loaded while the other is used.
// Vector views of the transposition buffer that may alias each other (the writes and the
// reads of one wave rely on their program order, with no fence in between).
typedef _Float16 f16x2_alias __attribute__((ext_vector_type(2), may_alias));
typedef _Float16 f16x8_alias __attribute__((ext_vector_type(8), may_alias));
struct Herm2Lds {
    union {
        struct {
            f16x8 xi[4][4][2][4][kVec];      // [part ch, cl, dh, dl][group][jj][g][n]: 16 lanes read 256 B
            _Float16 tr[4][4][32 * kTrRow];  // per wave, per part: [rho][pair][li][tile & 1]
        } loop;
        float2 sum[2][kHermRows][kVec];      // end of item: direct sums of waves 0 + 2, 1 + 3
    };
};

__global__ __launch_bounds__(256, 1) void herm_product_wave_kernel(const float2* __restrict__ H,
                                                                   const XFrag16* __restrict__ Xs,
                                                                   const XFrag16* __restrict__ Xt,
                                                                   const int2* __restrict__ items, int64_t N,
                                                                   float2* __restrict__ Pn, float2* __restrict__ Pt) {
    __shared__ Herm2Lds lds;
    const int2 bc = items[blockIdx.x];
    const int b = bc.x, c = bc.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int64_t rb0 = static_cast<int64_t>(b) * kHermRows;
    const int64_t cbeg = static_cast<int64_t>(c) * kHermChunk;
    const int64_t kend = min(N, cbeg + kHermChunk);
    const bool diag = cbeg <= rb0;
    const int64_t kbeg = diag ? rb0 : cbeg;
    const int nd = diag ? static_cast<int>((min(kend, rb0 + kHermRows) - rb0 + 31) / 32) : 0;
    const int64_t klow = diag ? rb0 + kHermRows : cbeg;
    const int nl = kend > klow ? static_cast<int>((kend - klow + 31) / 32) : 0;
    const int nsteps = nd + nl;
    const int mine = wave < nsteps ? (nsteps - wave + 3) / 4 : 0;  // steps of this wave

    // vectors of the row block in the i order of the transposed fragments
    {
        const XFrag16* src = Xt + static_cast<int64_t>(b) * 4 * 2 * kVec * 4;
        for (int e = threadIdx.x; e < 4 * 2 * kVec * 4; e += 256) {  // src order [group][jj][n][g]
            const XFrag16 x = src[e];
            const int g4 = e & 3, n = (e >> 2) % kVec, jj = (e / (4 * kVec)) & 1, G = e / (8 * kVec);
            lds.loop.xi[0][G][jj][g4][n] = x.ch;
            lds.loop.xi[1][G][jj][g4][n] = x.cl;
            lds.loop.xi[2][G][jj][g4][n] = x.dh;
            lds.loop.xi[3][G][jj][g4][n] = x.dl;
        }
    }
    __syncthreads();

    f32x4 acc[16][2][2];
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u][0] = acc[t][u][1] = f32x4{0, 0, 0, 0};
    f32x4 acct[2][2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < 2; ++u) acct[h][u][0] = acct[h][u][1] = f32x4{0, 0, 0, 0};
    // column i = rb0 + 16 t + li of tile t (rows past N clamped to N - 1: their outputs are
    // not stored and their vector entries are staged as zero)
    const int64_t i0 = rb0 + li;
    const float4* hbase = reinterpret_cast<const float4*>(H + (i0 < N ? i0 : N - 1) * N) + g;
    const int64_t tstride = 16 * N / 2;  // float4 per 16 columns
    auto kstep = [&](int st) -> int64_t { return st < nd ? kbeg + 32 * st : klow + 32 * (st - nd); };
    auto load = [&](float4 (&rb)[2][4], int j, int G) {
        // step j of this wave (clamped to its last; to the item's last when it has none)
        const int st = min(wave + 4 * min(j, max(mine - 1, 0)), nsteps - 1);
        const int64_t k0 = kstep(st);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int64_t i = rb0 + 16 * (2 * G + t) + li;
            const float4* col = (i < N) ? hbase + (2 * G + t) * tstride : hbase;
#pragma unroll
            for (int q = 0; q < 4; ++q) rb[t][q] = col[k0 / 2 + 4 * q];
        }
    };
    _Float16* trw = &lds.loop.tr[wave][0][0];
    // one group of two tiles of step st: direct products, and (below the diagonal block)
    // the transposed products through the wave's LDS buffer
    auto group = [&](const float4 (&rb)[2][4], const XFrag16 (&xb)[2], int G, bool low) {
        HFrag16 hf[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) hf[t] = presplit_h(rb[t]);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const HFrag16& h = hf[t];
            const f16x8 nbh = -h.bh, nbl = -h.bl;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const XFrag16& x = xb[u];
                f32x4 re = acc[2 * G + t][u][0], im = acc[2 * G + t][u][1];
                re = mma_f16(h.al, x.ch, re);
                re = mma_f16(h.ah, x.cl, re);
                re = mma_f16(h.ah, x.ch, re);
                re = mma_f16(h.bl, x.dh, re);
                re = mma_f16(h.bh, x.dl, re);
                re = mma_f16(h.bh, x.dh, re);
                im = mma_f16(h.al, x.dh, im);
                im = mma_f16(h.ah, x.dl, im);
                im = mma_f16(h.ah, x.dh, im);
                im = mma_f16(nbl, x.ch, im);
                im = mma_f16(nbh, x.cl, im);
                im = mma_f16(nbh, x.ch, im);
                acc[2 * G + t][u][0] = re;
                acc[2 * G + t][u][1] = im;
            }
        }
        if (!low) return;
        // the wave's LDS operations execute in order and these writes and reads address the
        // same buffer, so no fence separates them (the MFMAs around them may interleave)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
                const int rho = 8 * (e >> 1) + 4 * (e & 1) + g;
                _Float16* dst = trw + rho * kTrRow + 2 * li;
                const HFrag16& h0 = hf[0];
                const HFrag16& h1 = hf[1];
                *reinterpret_cast<f16x2_alias*>(dst) = f16x2{h0.ah[e], h1.ah[e]};
                *reinterpret_cast<f16x2_alias*>(dst + 32 * kTrRow) = f16x2{h0.al[e], h1.al[e]};
                *reinterpret_cast<f16x2_alias*>(dst + 64 * kTrRow) = f16x2{h0.bh[e], h1.bh[e]};
                *reinterpret_cast<f16x2_alias*>(dst + 96 * kTrRow) = f16x2{h0.bl[e], h1.bl[e]};
            }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
                const _Float16* src = trw + (16 * h + li) * kTrRow + 8 * g;
                const f16x8 ah = *reinterpret_cast<const f16x8_alias*>(src);
                const f16x8 al = *reinterpret_cast<const f16x8_alias*>(src + 32 * kTrRow);
                const f16x8 bh = *reinterpret_cast<const f16x8_alias*>(src + 64 * kTrRow);
                const f16x8 bl = *reinterpret_cast<const f16x8_alias*>(src + 96 * kTrRow);
                const f16x8 nbh = -bh, nbl = -bl;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    XFrag16 x;
                    x.ch = lds.loop.xi[0][G >> 1][G & 1][g][16 * u + li];
                    x.cl = lds.loop.xi[1][G >> 1][G & 1][g][16 * u + li];
                    x.dh = lds.loop.xi[2][G >> 1][G & 1][g][16 * u + li];
                    x.dl = lds.loop.xi[3][G >> 1][G & 1][g][16 * u + li];
                    f32x4 re = acct[h][u][0], im = acct[h][u][1];
                    re = mma_f16(al, x.ch, re);
                    re = mma_f16(ah, x.cl, re);
                    re = mma_f16(ah, x.ch, re);
                    re = mma_f16(nbl, x.dh, re);
                    re = mma_f16(nbh, x.dl, re);
                    re = mma_f16(nbh, x.dh, re);
                    im = mma_f16(al, x.dh, im);
                    im = mma_f16(ah, x.dl, im);
                    im = mma_f16(ah, x.dh, im);
                    im = mma_f16(bl, x.ch, im);
                    im = mma_f16(bh, x.cl, im);
                    im = mma_f16(bh, x.ch, im);
                    acct[h][u][0] = re;
                    acct[h][u][1] = im;
                }
            }
    };
    float4 bufa[2][4], bufb[2][4];
    load(bufa, 0, 0);
    load(bufb, 0, 1);
    for (int j = 0; j < mine; ++j) {
        const int st = wave + 4 * j;
        const int64_t k0 = kstep(st);
        const bool low = st >= nd;
        XFrag16 xb[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) xb[u] = Xs[((k0 / 32) * kVec + 16 * u + li) * 4 + g];
        // groups 0..7 alternate between the two register sets; each set is reloaded with the
        // group two ahead (in the next step for groups 6 and 7) as soon as it is consumed
        group(bufa, xb, 0, low);
        load(bufa, j, 2);
        group(bufb, xb, 1, low);
        load(bufb, j, 3);
        group(bufa, xb, 2, low);
        load(bufa, j, 4);
        group(bufb, xb, 3, low);
        load(bufb, j, 5);
        group(bufa, xb, 4, low);
        load(bufa, j, 6);
        group(bufb, xb, 5, low);
        load(bufb, j, 7);
        group(bufa, xb, 6, low);
        load(bufa, j + 1, 0);
        group(bufb, xb, 7, low);
        load(bufb, j + 1, 1);
        if (low) {
            float2* out = Pt + static_cast<int64_t>(b) * N * kVec;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int64_t k = k0 + 16 * h + 8 * (g >> 1) + (g & 1) + 2 * v;
                        if (k < N) out[k * kVec + 16 * u + li] = make_float2(acct[h][u][0][v], acct[h][u][1][v]);
                        acct[h][u][0][v] = 0.0f;
                        acct[h][u][1][v] = 0.0f;
                    }
        }
    }
    // direct sums of the four waves: waves 0, 1 store, waves 2, 3 add, all write Pn
    __syncthreads();  // every wave is done with the loop's LDS
    auto deposit = [&](bool add) {
        float2(&slot)[kHermRows][kVec] = lds.sum[wave & 1];
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float2& d = slot[16 * t + 4 * g + v][16 * u + li];
                    const float2 x = make_float2(acc[t][u][0][v], acc[t][u][1][v]);
                    d = add ? make_float2(d.x + x.x, d.y + x.y) : x;
                }
    };
    if (wave < 2) deposit(false);
    __syncthreads();
    if (wave >= 2) deposit(true);
    __syncthreads();
    float2* out = Pn + static_cast<int64_t>(c) * N * kVec;
    for (int e = threadIdx.x; e < kHermRows * kVec; e += 256) {
        const int r = e / kVec, n = e % kVec;
        const int64_t i = rb0 + r;
        if (i < N) {
            const float2 a = lds.sum[0][r][n], d = lds.sum[1][r][n];
            out[i * kVec + n] = make_float2(a.x + d.x, a.y + d.y);
        }
    }
}

// Y[r][n] (column-major, FP64) = scale[n] (sum over the items' direct partial sums of r and
// the transposed partial sums of r from the row blocks above it); 64 rows per workgroup
// (all in one row block), fixed summation order.
constexpr int kReduceRows = 8;  // rows per workgroup of the reduction (256 threads = 8 rows x 32 vectors)
__global__ __launch_bounds__(256) void herm_reduce_kernel(const float2* __restrict__ Pn,
                                                          const float2* __restrict__ Pt, int64_t N, int nvec,
                                                          int nchunks, const double* __restrict__ scale,
                                                          double2* __restrict__ Y) {
    constexpr int kRows = kReduceRows;  // one output (row, vector) per thread
    __shared__ double2 tile[kRows][kVec + 1];
    // the last rows (the most transposed-block slots) first
    const int64_t r0 = static_cast<int64_t>(gridDim.x - 1 - blockIdx.x) * kRows;
    const int b = static_cast<int>(r0 / kHermRows);
    const int c0 = static_cast<int>(static_cast<int64_t>(b) * kHermRows / kHermChunk);
#pragma unroll
    for (int j = 0; j < kRows * kVec / 256; ++j) {
        const int f = threadIdx.x + 256 * j;
        const int rr = f / kVec, n = f % kVec;
        const int64_t r = r0 + rr;
        double sr = 0.0, si = 0.0;
        if (r < N && n < nvec) {
            // terms: the direct partial sums of chunks c0.., then the transposed ones of row
            // blocks 0..b-1, summed in that order in eight interleaved partial sums
            const int nc = nchunks - c0, nt = nc + b;
            const int64_t off = r * kVec + n, stride = N * kVec;
            const float2* pn = Pn + static_cast<int64_t>(c0) * stride + off;
            const float2* pt = Pt + off;
            double ar[8] = {0, 0, 0, 0, 0, 0, 0, 0}, ai[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int t = 0;
            for (; t + 8 <= nt; t += 8) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int u = t + j;
                    const float2 v = u < nc ? pn[u * stride] : pt[(u - nc) * stride];
                    ar[j] += v.x;
                    ai[j] += v.y;
                }
            }
            for (int j = 0; t < nt; ++t, ++j) {
                const float2 v = t < nc ? pn[t * stride] : pt[(t - nc) * stride];
                ar[j] += v.x;
                ai[j] += v.y;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                sr += ar[j];
                si += ai[j];
            }
            sr *= scale[n];
            si *= scale[n];
        }
        tile[rr][n] = make_double2(sr, si);
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

// Back from the split layout to FP32 complex entries, in place: a wave converts four
// consecutive 256-byte column steps (128 entries; N is a multiple of 32, so the steps of a
// column are consecutive in memory) per pass, entry = (hi + lo) / scale for both parts.
// Steps of the block-upper part (never built) are rewritten the same way; the mirror of
// the block-lower part overwrites them afterwards.
__global__ __launch_bounds__(256) void unsplit_kernel(float2* __restrict__ H, int64_t N, float inv_scale) {
    __shared__ f16x8 piece[4][64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t nsteps = (N / 32) * N;  // 32-entry steps of the whole matrix, in memory order
    for (int64_t s0 = (static_cast<int64_t>(blockIdx.x) * 4 + w) * 4; s0 < nsteps;
         s0 += static_cast<int64_t>(gridDim.x) * 16) {
        const int ns = static_cast<int>(min<int64_t>(4, nsteps - s0));
        float2* base = H + 32 * s0;
        if (l < 16 * ns) piece[w][l] = reinterpret_cast<const f16x8*>(base)[l];
        wave_lds_sync();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int o = l + 64 * h;  // entry o of the four steps: step o / 32, row o % 32
            const int st = o >> 5, row = o & 31;
            if (st < ns) {
                // row = 8 (e >> 1) + 2 g + (e & 1)
                const int g = (row >> 1) & 3, e = 2 * (row >> 3) + (row & 1);
                const f16x8* p = piece[w] + 16 * st;
                const float a = static_cast<float>(p[g][e]) + static_cast<float>(p[4 + g][e]);
                const float b = static_cast<float>(p[8 + g][e]) + static_cast<float>(p[12 + g][e]);
                base[o] = make_float2(a * inv_scale, b * inv_scale);
            }
        }
        wave_lds_sync();
    }
}

// Zeroes rows 0..31 of the columns i >= kSplitRowBlock, which a block-lower build leaves
// unwritten: the last step of a column reads up to 31 entries into the next column, where
// they meet zero vector entries and must be finite.
__global__ void zero_heads_kernel(float2* __restrict__ H, int64_t N) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= (N - kSplitRowBlock) * 32) return;
    const int64_t i = kSplitRowBlock + idx / 32;
    H[i * N + idx % 32] = make_float2(0.0f, 0.0f);
}

// Fills H(k, i) = conj(H(i, k)) for the rows k < kSplitRowBlock floor(i / kSplitRowBlock) of
// column i (the block-upper part left out by a block-lower build), 64 x 64 tiles through
// LDS; grid (tiles, tiles), block 256, tiles entirely outside that part exit.
__global__ __launch_bounds__(256) void mirror_upper_kernel(float2* __restrict__ H, int64_t N) {
    __shared__ float2 tile[64][65];
    const int64_t k0 = static_cast<int64_t>(blockIdx.x) * 64, i0 = static_cast<int64_t>(blockIdx.y) * 64;
    const int64_t thr0 = i0 / kSplitRowBlock * kSplitRowBlock;
    if (k0 >= thr0) return;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // source: H(i, k) for i in the column tile, k in the row tile, read down column k
    for (int r = ty; r < 64; r += 4) {
        const int64_t k = k0 + r, i = i0 + tx;
        if (k < N && i < N) tile[r][tx] = H[k * N + i];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int64_t i = i0 + r, k = k0 + tx;
        if (i < N && k < N && k < i / kSplitRowBlock * kSplitRowBlock) {
            const float2 v = tile[tx][r];
            H[i * N + k] = make_float2(v.x, -v.y);
        }
    }
}

}  // namespace

static void lay_out_items(Context& c);

void split_product_scale(Context& c) {
    const size_t nft = static_cast<size_t>(c.lat.natoms) * c.lat.natoms * c.lat.ntab;
    const int nftq = c.lat.natoms * c.lat.natoms;
    const int64_t ncols = static_cast<int64_t>(c.nk) * c.nb;
    constexpr int kBlocks = 64;
    DevBuf<double>& part = c.s_bounds;
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
    if (c.h16_scale > 0.0f) lay_out_items(c);
}

// Work items (row block, column chunk) of the split product, the longest first; laid out
// once per matrix dimension.
static void lay_out_items(Context& c) {
    const int64_t N = c.N;
    if (c.herm_n == N && c.herm_items > 0) return;
    const int B = static_cast<int>((N + kHermRows - 1) / kHermRows);
    const int C = static_cast<int>((N + kHermChunk - 1) / kHermChunk);
    std::vector<int2> it;
    std::vector<int64_t> len;
    for (int ch = 0; ch < C; ++ch)
        for (int b = 0; b < B; ++b) {
            const int64_t rb0 = static_cast<int64_t>(b) * kHermRows, cb = static_cast<int64_t>(ch) * kHermChunk;
            if (rb0 >= cb + kHermChunk) continue;
            it.push_back(make_int2(b, ch));
            len.push_back(std::min<int64_t>(N, cb + kHermChunk) - std::max(rb0, cb));
        }
    std::vector<size_t> order(it.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return len[x] > len[y]; });
    std::vector<int2> sorted(it.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = it[order[i]];
    c.d_herm_items.alloc(sorted.size());
    XATU_HIP(hipMemcpy(c.d_herm_items.p, sorted.data(), sorted.size() * sizeof(int2), hipMemcpyHostToDevice));
    c.herm_items = static_cast<int>(sorted.size());
    c.herm_n = N;
}

void clear_upper_heads(Context& c) {
    if (c.N <= kSplitRowBlock) return;
    const int64_t n = (c.N - kSplitRowBlock) * 32;
    hipLaunchKernelGGL(zero_heads_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, c.d_H32.p, c.N);
    XATU_HIP(hipGetLastError());
}

void complete_matrix32(Context& c) {
    if (!c.built32 || !c.h32_lower) return;
    if (c.h32_presplit) {
        hipLaunchKernelGGL(unsplit_kernel, dim3(2048), dim3(256), 0, c.stream, c.d_H32.p, c.N, 1.0f / c.h16_scale);
        XATU_HIP(hipGetLastError());
        c.h32_presplit = false;
    }
    const unsigned t = static_cast<unsigned>((c.N + 63) / 64);
    hipLaunchKernelGGL(mirror_upper_kernel, dim3(t, t), dim3(256), 0, c.stream, c.d_H32.p, c.N);
    XATU_HIP(hipGetLastError());
    c.h32_lower = false;
}

bool split_layout_available(const Context& c) {
    return c.h16_scale > 0.0f && (c.N % 2) == 0 && c.herm_n == c.N;
}

bool split_product_available(const Context& c) {
    return c.built32 && !c.fp32_arith_only && split_layout_available(c);
}

void launch_matvec_split16(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    const int64_t N = c.N;
    const int B = static_cast<int>((N + kHermRows - 1) / kHermRows);
    const int C = static_cast<int>((N + kHermChunk - 1) / kHermChunk);
    const int64_t nsteps = (N + 31) / 32;
    const int64_t ngroups = 4 * static_cast<int64_t>(B);
    const int64_t rows = (N + kMaxChunks - 1) / kMaxChunks;
    const int chunks = static_cast<int>((N + rows - 1) / rows);
    const size_t xs_n = static_cast<size_t>(nsteps) * kVec * 4 * sizeof(XFrag16) / sizeof(float4);
    const size_t xt_n = static_cast<size_t>(ngroups) * 2 * kVec * 4 * sizeof(XFrag16) / sizeof(float4);
    c.s_xs16.alloc(std::max(c.s_xs16.n, xs_n));
    c.s_xt16.alloc(std::max(c.s_xt16.n, xt_n));
    c.s_yp32.alloc(std::max<size_t>(c.s_yp32.n, static_cast<size_t>(C) * N * kVec));
    c.s_pt32.alloc(std::max<size_t>(c.s_pt32.n, static_cast<size_t>(std::max(1, B - 1)) * N * kVec));
    c.s_colmax.alloc(static_cast<size_t>(kVec) * kMaxChunks);
    c.s_xscale.alloc(kVec);
    c.s_xinv.alloc(kVec);
    XFrag16* xs = reinterpret_cast<XFrag16*>(c.s_xs16.p);
    XFrag16* xt = reinterpret_cast<XFrag16*>(c.s_xt16.p);
    const float sH = c.h16_scale;
    const int64_t nstage = (nsteps + 2 * ngroups) * kVec * 4;
    for (int v0 = 0; v0 < nvec; v0 += kVec) {
        const int nv = std::min(kVec, nvec - v0);
        const double2* Xb = X + static_cast<int64_t>(v0) * N;
        hipLaunchKernelGGL(column_max_kernel, dim3(chunks, kVec), dim3(256), 0, s, Xb, N, nv, rows, c.s_colmax.p);
        hipLaunchKernelGGL(vector_scale_kernel, dim3(1), dim3(64), 0, s, c.s_colmax.p, chunks, sH, c.s_xscale.p,
                           c.s_xinv.p);
        hipLaunchKernelGGL(stage_split_kernel, dim3((nstage + 255) / 256), dim3(256), 0, s, Xb, N, nv, nsteps, ngroups,
                           c.s_xscale.p, xs, xt);
        if (c.h32_presplit) {
            hipLaunchKernelGGL(herm_product_wave_kernel, dim3(c.herm_items), dim3(256), 0, s, c.d_H32.p, xs, xt,
                               c.d_herm_items.p, N, c.s_yp32.p, c.s_pt32.p);
        } else {
            hipLaunchKernelGGL((herm_product_kernel<kHermTiles, false>), dim3(c.herm_items), dim3(256), 0, s,
                               c.d_H32.p, xs, xt, c.d_herm_items.p, N, sH, c.s_yp32.p, c.s_pt32.p);
        }
        hipLaunchKernelGGL(herm_reduce_kernel, dim3((N + kReduceRows - 1) / kReduceRows), dim3(256), 0, s, c.s_yp32.p, c.s_pt32.p, N, nv, C,
                           c.s_xinv.p, Y + static_cast<int64_t>(v0) * N);
    }
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
