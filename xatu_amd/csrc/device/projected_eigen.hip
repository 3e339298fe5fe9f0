// Rayleigh-Ritz eigensolver of the Davidson projected matrix, on the device: the lowest
// nwant eigenpairs of a Hermitian m x m matrix (m <= 128, nwant <= 32) in one workgroup
// of 1024 threads, so that the solver's iterations need no host round trip for it.
//
//  1. Householder reduction to a Hermitian tridiagonal, the same reflections as the host
//     solver (common/heev.hpp): column k below the diagonal, x, goes to alpha e_1 with
//     alpha = -phase(x_0) |x| by I - 2 v v^H, v = (x - alpha e_1) / |x - alpha e_1|, and the
//     matrix becomes (I - 2 v v^H) A (I - 2 v v^H) = A - 2 (v w^H + w v^H), p = A v,
//     w = p - (v^H p) v. The matrix lives in registers: wave w holds rows 8 w .. 8 w + 7, lane l
//     row 8 w + l / 8 and columns l mod 8 + 8 j (j < 16). Step k: p = A v_k for the wave's
//     rows (summed over the 8 lanes of a row) and v_k^H p over the wave [one barrier]; then
//     every wave derives column k + 1 of the updated matrix itself (from row k + 1 as it was
//     before the update, kept in LDS, and v_k, p) and forms v_{k+1} into its own LDS copy,
//     and applies the rank-2 update of step k. These two are independent: even waves do them
//     in one order, odd waves in the other, so that one's latency overlaps the other's
//     arithmetic. Waves whose rows are all eliminated (or beyond m) drop out, and columns at
//     or left of k are skipped.
//  2. A diagonal phase change makes the tridiagonal real (T_complex = Ph T Ph^H); the lowest
//     nwant eigenvalues by multisection on Sturm counts (1024 / nwant lanes per eigenvalue,
//     pivmin-guarded pivots as in LAPACK dstebz; the pivot reciprocal by the hardware
//     estimate and one Newton step).
//  3. Their eigenvectors by inverse iteration, one lane each (LU with partial pivoting of
//     T - lambda I, refactored with every forward solve so that only U is stored; from a
//     fixed pseudo-random start until the growth test of LAPACK dstein passes, then once
//     more, at most four solves; the host solver's shifts for coincident eigenvalues);
//     each vector is then orthogonalised against the earlier vectors of its cluster of
//     close eigenvalues (modified Gram-Schmidt).
//  4. Back-transformation y = H_0 ... H_{m-3} Ph z, 16 lanes per vector, the reflectors
//     staged through LDS.
#include <cfloat>

#include "context.hpp"
#include "cplx.hpp"

namespace xatu_dev {
namespace {

constexpr int kEM = 128;         // largest projected dimension
constexpr int kEThreads = 1024;  // one workgroup
constexpr int kEGroups = 8;      // column groups of the register layout
constexpr int kECols = 16;       // columns per thread
constexpr int kEWant = 32;       // most eigenpairs
constexpr int kEPad = kEM + 1;   // padded row of the per-eigenvector arrays
constexpr int kEStage = 32;      // reflectors per LDS stage of the back-transformation

template <typename TR>
struct EigLds {
    union {
        struct {
            cx_t<TR> v[2][16][kEM];       // per-wave copy of the reflector, by step parity
            cx_t<TR> p[2][kEM];           // p of the step, by step parity
            cx_t<TR> row[2][kEM];         // row k + 1 before the update of step k, by parity of k + 1
            TR cpart[2][16];              // v^H p over the rows of each wave, by step parity
        } tri;
        struct {
            double ui[kEWant][kEPad], u1[kEWant][kEPad], u2[kEWant][kEPad];  // U rows; ui = 1 / u0
        } inv;
        double2 refl[kEStage][kEM];
    };
    double z[kEWant][kEPad];  // real tridiagonal eigenvectors
    double d[kEM], e[kEM], e2[kEM];
    double2 ec[kEM], ph[kEM];
    double lam[kEWant], shift[kEWant];
    int cstart[kEWant];
    double tnorm, pivmin, glo, ghi;
};

// Cross-lane moves of a double by DPP (two 32-bit halves); lanes whose row is outside
// row_mask receive 0.
template <int kCtrl, int kRowMask = 0xF>
__device__ inline double dpp(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), kCtrl, kRowMask, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), kCtrl, kRowMask, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <int kCtrl, int kRowMask = 0xF>
__device__ inline float dpp(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, kRowMask, 0xF, false));
}
__device__ inline double lane63(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), 63);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), 63);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ inline float lane63(float x) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63)); }
constexpr int kDppXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141; // lane i <-> 7 - i within 8
constexpr int kDppMirror = 0x140;     // lane i <-> 15 - i within 16
constexpr int kDppBcast15 = 0x142;    // lane 15 of a row to the next row
constexpr int kDppBcast31 = 0x143;    // lane 31 to rows 2 and 3

// Sum over the 8 lanes l ^ 0..7 (every lane of the group receives it).
template <typename T>
__device__ inline T sum8(T x) {
    x += dpp<kDppXor1>(x);
    x += dpp<kDppXor2>(x);
    x += dpp<kDppHalfMirror>(x);
    return x;
}
// Sum over the 16 lanes of a DPP row (every lane of the row receives it).
template <typename T>
__device__ inline T sum16_row(T x) {
    x = sum8(x);
    x += dpp<kDppMirror>(x);
    return x;
}
// Sum over the wave of a value that is the same within groups of 8 lanes: one term per group.
template <typename T>
__device__ inline T wave_sum_groups8(T x) {
    x += dpp<kDppMirror>(x);             // two groups, every lane of a row of 16
    x += dpp<kDppBcast15, 0xA>(x);       // rows 1, 3 += rows 0, 2
    x += dpp<kDppBcast31, 0xC>(x);       // rows 2, 3 += row 1
    return lane63(x);
}
// Sum over the wave, the same in every lane.
template <typename T>
__device__ inline T wave_sum(T x) { return wave_sum_groups8(sum8(x)); }

__device__ inline double2 csc(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline float2 csc(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ inline float rsq(float x) { return rsqrtf(x); }
__device__ inline double rsq(double x) { return rsqrt(x); }

// Pairwise sum of 16 values (one term per wave).
template <typename T>
__device__ inline T sum16(const T* a) {
    T t[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] = a[i] + a[i + 8];
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] += t[i + 4];
    return (t[0] + t[2]) + (t[1] + t[3]);
}

__device__ inline void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Workgroup barrier that waits for this wave's LDS operations only (its global stores,
// the reflectors read after a full __syncthreads, stay in flight).
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// 1 / q for normal q: the hardware estimate refined by one Newton step.
__device__ inline double recip(double q) {
    double y = __builtin_amdgcn_rcp(q);
    const double e = fma(-q, y, 1.0);
    return fma(y, e, y);
}

// TR: arithmetic of the tridiagonalisation (float or double); the eigenvalues, eigenvectors
// and back-transformation are FP64.
template <typename TR>
__global__ __launch_bounds__(kEThreads) void projected_eigen_kernel(const double2* __restrict__ A, int lda, int m,
                                                                    int nwant, cx_t<TR>* __restrict__ refl,
                                                                    double* __restrict__ theta,
                                                                    double2* __restrict__ Y, int ldy) {
    using CT = cx_t<TR>;
    // working accuracy of the tridiagonal: the eigenvalues are resolved to it (multisection
    // tolerance, shift separation, inverse-iteration growth test), not beyond
    constexpr double kTolEps = sizeof(TR) == 4 ? static_cast<double>(FLT_EPSILON) : DBL_EPSILON;
    __shared__ EigLds<TR> s;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int r = 8 * w + (lane >> 3), cg = lane & 7;
    const int duty = (m - 1) >> 3;  // the wave of row m - 1 stores the reflectors and the tridiagonal

    // ---- 1. tridiagonalisation -------------------------------------------------------
    CT B[kECols];  // A[r][cg + 8 j], Hermitian from the lower triangle
#pragma unroll
    for (int j = 0; j < kECols; ++j) {
        const int c = cg + kEGroups * j;
        double2 a = make_double2(0.0, 0.0);
        if (r < m && c < m) a = (r >= c) ? A[r + static_cast<size_t>(c) * lda] : cconj(A[c + static_cast<size_t>(r) * lda]);
        if (r == c) a.y = TR(0);
        B[j] = mk(static_cast<TR>(a.x), static_cast<TR>(a.y));
    }
    // the state of a step -1 that changes nothing
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        s.tri.v[b][w][lane] = mk(TR(0), TR(0));
        s.tri.v[b][w][lane + 64] = mk(TR(0), TR(0));
    }
    if (t < 2 * kEM) s.tri.p[t >> 7][t & (kEM - 1)] = mk(TR(0), TR(0));
    if (t < 32) s.tri.cpart[t >> 4][t & 15] = TR(0);
    if (r == 0) {
#pragma unroll
        for (int j = 0; j < kECols; ++j) s.tri.row[0][cg + kEGroups * j] = B[j];
    }
    __syncthreads();

    const int jhi = (m + 7) >> 3;  // columns cg + 8 j, j >= jhi, are all >= m
    auto alive_at = [&](int k) { return 8 * w + 7 > k && 8 * w < m; };  // rows in (k, m)
    // v_k into this wave's copy: column k of the matrix after the update of step k - 1, from
    // row k before that update (row buffer) and v_{k-1}, p_{k-1}
    auto reflector = [&](int k) {
        const int cur = k & 1, prv = cur ^ 1;
        const TR cprev = sum16(s.tri.cpart[prv]);
        const CT* vp = s.tri.v[prv][w];
        const CT vpk = vp[k];
        const CT wpk = csub(s.tri.p[prv][k], csc(vpk, cprev));
        auto row_k = [&](int i) -> CT {
            const CT vpi = vp[i];
            const CT wpi = csub(s.tri.p[prv][i], csc(vpi, cprev));
            CT a = s.tri.row[cur][i];
            a = csub(a, csc(cmul(vpk, cconj(wpi)), TR(2)));
            a = csub(a, csc(cmul(wpk, cconj(vpi)), TR(2)));
            return a;
        };
        const TR dk = row_k(k).x;
        const CT x0 = cconj(row_k(k + 1));
        const CT xa = cconj(row_k(lane)), xb = cconj(row_k(lane + 64));  // x_i = conj(A[k][i])
        const bool ina = lane > k + 1 && lane < m, inb = lane + 64 > k + 1 && lane + 64 < m;
        TR tl = TR(0);
        if (ina) tl += xa.x * xa.x + xa.y * xa.y;
        if (inb) tl += xb.x * xb.x + xb.y * xb.y;
        tl = wave_sum(tl);  // |x|^2 without x_0
        const TR a0 = x0.x * x0.x + x0.y * x0.y;
        const TR xn2 = tl + a0;
        CT va = mk(TR(0), TR(0)), vb = mk(TR(0), TR(0)), ek = x0;
        if (!(xn2 == TR(0) || tl <= TR(0))) {
            // alpha = -x0 |x| / |x0| (= -|x| for x0 = 0), u0 = x0 - alpha
            CT alpha, u0;
            if (a0 > TR(0)) {
                const TR f = sqrt(xn2 / a0);
                alpha = csc(x0, -f);
                u0 = csc(x0, TR(1) + f);
            } else {
                alpha = mk(-sqrt(xn2), TR(0));
                u0 = mk(-alpha.x, TR(0));
            }
            const TR inv = rsq(tl + u0.x * u0.x + u0.y * u0.y);
            if (ina) va = csc(xa, inv);
            if (inb) vb = csc(xb, inv);
            if (lane == k + 1) va = csc(u0, inv);
            if (lane + 64 == k + 1) vb = csc(u0, inv);
            ek = alpha;
        }
        s.tri.v[cur][w][lane] = va;
        s.tri.v[cur][w][lane + 64] = vb;
        if (w == duty) {
            refl[static_cast<size_t>(k) * kEM + lane] = va;
            refl[static_cast<size_t>(k) * kEM + lane + 64] = vb;
            if (lane == 0) {
                s.d[k] = dk;
                s.ec[k] = make_double2(ek.x, ek.y);
            }
        }
    };
    // A -= 2 (v w^H + w v^H) of step k on the columns right of k
    auto update = [&](int k) {
        const int cur = k & 1;
        const int jlo = (k + 1) >> 3;  // columns cg + 8 j, j < jlo, are all <= k
        const TR cc = sum16(s.tri.cpart[cur]);
        const CT* vk = s.tri.v[cur][w];
        const CT vr = vk[r];
        const CT wr = csub(s.tri.p[cur][r], csc(vr, cc));
        const CT v2 = csc(vr, TR(2)), w2 = csc(wr, TR(2));
#pragma unroll
        for (int j = 0; j < kECols; ++j) {
            const int c = cg + kEGroups * j;
            if (j >= jlo && j < jhi && c > k) {
                const CT vc = vk[c];
                const CT pc = s.tri.p[cur][c];
                const CT wc = mk(fma(-cc, vc.x, pc.x), fma(-cc, vc.y, pc.y));
                // a -= 2 v_r conj(w_c) + 2 w_r conj(v_c)
                CT a = B[j];
                a.x = fma(-v2.x, wc.x, a.x);
                a.x = fma(-v2.y, wc.y, a.x);
                a.x = fma(-w2.x, vc.x, a.x);
                a.x = fma(-w2.y, vc.y, a.x);
                a.y = fma(v2.x, wc.y, a.y);
                a.y = fma(-v2.y, wc.x, a.y);
                a.y = fma(w2.x, vc.y, a.y);
                a.y = fma(-w2.y, vc.x, a.y);
                if (r == c) a.y = TR(0);
                B[j] = a;
            }
        }
    };

    if (m >= 3 && alive_at(0)) {
        reflector(0);
        wave_lds_fence();
    }
    for (int k = 0; k + 2 < m; ++k) {
        const int cur = k & 1;
        const bool alive = alive_at(k);
        const int jlo = (k + 1) >> 3;
        if (alive) {
            if (r == k + 1) {  // row k + 1 before the update of step k
#pragma unroll
                for (int j = 0; j < kECols; ++j) s.tri.row[(k + 1) & 1][cg + kEGroups * j] = B[j];
            }
            // p = A v_k for the rows of this wave: over the thread's columns, then the row's 8 lanes
            const CT* vk = s.tri.v[cur][w];
            CT acc = mk(TR(0), TR(0)), acc2 = mk(TR(0), TR(0));
#pragma unroll
            for (int j = 0; j < kECols; j += 2) {
                if (j >= jlo && j < jhi) cfma(acc, B[j], vk[cg + kEGroups * j]);
                if (j + 1 >= jlo && j + 1 < jhi) cfma(acc2, B[j + 1], vk[cg + kEGroups * (j + 1)]);
            }
            acc.x = sum8(acc.x + acc2.x);
            acc.y = sum8(acc.y + acc2.y);
            if (cg == 0) s.tri.p[cur][r] = acc;
            const CT vr = vk[r];
            // Re(conj(v_r) p_r), the same in the row's 8 lanes, summed over the wave's rows
            const TR cs = wave_sum_groups8(vr.x * acc.x + vr.y * acc.y);
            if (lane == 0) s.tri.cpart[cur][w] = cs;
        } else if (lane == 0) {
            s.tri.cpart[cur][w] = TR(0);
        }
        lds_barrier();
        const bool next = k + 3 < m && alive_at(k + 1);
        if (w & 1) {
            if (alive) update(k);
            if (next) reflector(k + 1);
        } else {
            if (next) reflector(k + 1);
            if (alive) update(k);
        }
        wave_lds_fence();
    }
    // the last 2 x 2 block
#pragma unroll
    for (int j = 0; j < kECols; ++j) {
        const int c = cg + kEGroups * j;
        if (m >= 2) {
            if (r == m - 2 && c == m - 2) s.d[m - 2] = B[j].x;
            if (r == m - 1 && c == m - 2) s.ec[m - 2] = make_double2(B[j].x, B[j].y);
        }
        if (r == m - 1 && c == m - 1) s.d[m - 1] = B[j].x;
    }
    __syncthreads();

    // ---- 2. real tridiagonal, bounds; lowest eigenvalues -------------------------------
    if (t == 0) {
        double2 ph = make_double2(1.0, 0.0);
        s.ph[0] = ph;
        double e2max = 0.0;
        for (int k = 0; k + 1 < m; ++k) {
            const double ae = hypot(s.ec[k].x, s.ec[k].y);
            s.e[k] = ae;
            s.e2[k] = ae * ae;
            e2max = fmax(e2max, ae * ae);
            if (ae > 0.0) ph = cmul(ph, csc(s.ec[k], 1.0 / ae));
            s.ph[k + 1] = ph;
        }
        s.e[m - 1] = 0.0;
        s.e2[m - 1] = 0.0;
        double tn = 0.0, lo = s.d[0], hi = s.d[0];
        for (int i = 0; i < m; ++i) {
            const double off = s.e[i] + (i ? s.e[i - 1] : 0.0);
            tn = fmax(tn, fabs(s.d[i]) + off);
            lo = fmin(lo, s.d[i] - off);
            hi = fmax(hi, s.d[i] + off);
        }
        s.tnorm = tn;
        s.pivmin = DBL_MIN * fmax(1.0, e2max);
        const double fudge = 1e-12 * tn + 4.0 * s.pivmin;
        s.glo = lo - fudge;
        s.ghi = hi + fudge;
    }
    __syncthreads();
    {
        int nwp = 1;
        while (nwp < nwant) nwp <<= 1;
        // lanes per eigenvalue, a power of two: 256 lanes in all (fewer Sturm sequences in
        // all than with every lane busy, as the extra iterations cost less than the lanes)
        const int G = min(64, 256 / nwp);
        const int ev = t / G, gs = t % G;
        const double pivmin = s.pivmin, tnorm = s.tnorm;
        if (ev < nwant) {
            double lo = s.glo, hi = s.ghi;
            const int gshift = lane - gs;  // first lane of the group in the wave
            const unsigned long long gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
            for (int it = 0; it < 64; ++it) {
                if (hi - lo <= kTolEps * (2.0 * fmax(fabs(lo), fabs(hi)) + tnorm)) break;
                const double step = (hi - lo) / (G + 1);
                const double x = lo + step * (gs + 1);
                // Sturm count: eigenvalues below x
                int cnt = 0;
                double q = s.d[0] - x;
                if (fabs(q) < pivmin) q = -pivmin;
                if (q <= 0.0) ++cnt;
                for (int i = 1; i < m; ++i) {
                    q = (s.d[i] - x) - s.e2[i - 1] * recip(q);
                    if (fabs(q) < pivmin) q = -pivmin;
                    if (q <= 0.0) ++cnt;
                }
                const unsigned long long bal = __ballot(cnt >= ev + 1);
                const unsigned long long mine = (bal >> gshift) & gmask;
                const int first = mine ? __ffsll(static_cast<long long>(mine)) - 1 : G;
                const double nlo = first == 0 ? lo : lo + step * first;
                const double nhi = first == G ? hi : lo + step * (first + 1);
                lo = nlo;
                hi = nhi;
            }
            if (gs == 0) s.lam[ev] = 0.5 * (lo + hi);
        }
    }
    __syncthreads();
    if (t == 0) {
        // clusters of close eigenvalues and separated shifts for coincident ones
        const double cluster = 1e-3 * fmax(s.tnorm, 1e-300);
        const double sep = 10.0 * kTolEps * s.tnorm;
        int cs = 0;
        for (int i = 0; i < nwant; ++i) {
            if (i > 0 && s.lam[i] - s.lam[i - 1] > cluster) cs = i;
            s.cstart[i] = cs;
            double sh = s.lam[i];
            if (i > cs && sh - s.shift[i - 1] < sep) sh = s.shift[i - 1] + sep;
            s.shift[i] = sh;
        }
    }
    __syncthreads();

    // ---- 3. inverse iteration, one lane per eigenvector --------------------------------
    // (the recurrences carry their running values in registers, four rows per pass with the
    // pass's loads ahead of its stores; LDS holds U and z)
    if (t < nwant) {
        double* ui = s.inv.ui[t];
        double* u1 = s.inv.u1[t];
        double* u2 = s.inv.u2[t];
        double* z = s.z[t];
        const double lam = s.shift[t];
        const double eps = DBL_EPSILON * fmax(s.tnorm, DBL_MIN);
        auto guard = [&](double x) { return (fabs(x) < eps) ? (x < 0.0 ? -eps : eps) : x; };
        unsigned sd = static_cast<unsigned>(t + 1) * 2654435761u + 12345u;
        for (int i = 0; i < m; ++i) {
            sd = sd * 1664525u + 1013904223u;
            z[i] = 0.5 + (sd >> 8) * (1.0 / 16777216.0);
        }
        double scale = 1.0;  // the previous solve's normalisation, applied on the way in
        // a solve whose growth |(T - lam I)^-1 z| reaches 1 / (64 sqrt(m) eps |T|), eps the
        // working accuracy, has converged (the test of LAPACK dstein); one more solve follows
        // it, at most four
        const double grow = 1.0 / (64.0 * sqrt(static_cast<double>(m)) * kTolEps * fmax(s.tnorm, DBL_MIN));
        int extra = -1;
        for (int it = 0; it < 4 && extra != 0; ++it) {
            // LU with partial pivoting of T - lam I (pending row: alpha at i, beta at i + 1)
            // and L y = z by the same row operations
            double alpha = s.d[0] - lam, beta = m > 1 ? s.e[0] : 0.0;
            double cur = z[0] * scale;
            for (int i0 = 0; i0 + 1 < m; i0 += 4) {
                double nz[4], sb[4], nd[4], ns[4], out[4], oi[4], o1[4], o2[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = min(i0 + u, m - 2);
                    nz[u] = z[i + 1];
                    sb[u] = s.e[i];
                    nd[u] = s.d[i + 1] - lam;
                    ns[u] = (i + 2 < m) ? s.e[i + 1] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (i0 + u + 1 < m) {
                        double nxt = nz[u] * scale;
                        if (fabs(alpha) >= fabs(sb[u])) {
                            alpha = guard(alpha);
                            const double ia = 1.0 / alpha;
                            const double mult = sb[u] * ia;
                            oi[u] = ia;
                            o1[u] = beta;
                            o2[u] = 0.0;
                            alpha = nd[u] - mult * beta;
                            beta = ns[u];
                            nxt -= mult * cur;
                        } else {  // rows i and i + 1 exchanged
                            const double is = 1.0 / sb[u];
                            const double mult = alpha * is;
                            oi[u] = is;
                            o1[u] = nd[u];
                            o2[u] = ns[u];
                            alpha = beta - mult * nd[u];
                            beta = -mult * ns[u];
                            const double a = cur;
                            cur = nxt;
                            nxt = a - mult * cur;
                        }
                        out[u] = cur;
                        cur = nxt;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (i0 + u + 1 < m) {
                        z[i0 + u] = out[u];
                        ui[i0 + u] = oi[u];
                        u1[i0 + u] = o1[u];
                        u2[i0 + u] = o2[u];
                    }
                }
            }
            z[m - 1] = cur;
            ui[m - 1] = 1.0 / guard(alpha);
            u1[m - 1] = 0.0;
            u2[m - 1] = 0.0;
            // U x = y, the two following entries in registers
            double z1 = 0.0, z2 = 0.0, nrm = 0.0;
            for (int i0 = m - 1; i0 >= 0; i0 -= 4) {
                double yy[4], a1[4], a2[4], ia[4], out[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = max(i0 - u, 0);
                    yy[u] = z[i];
                    a1[u] = u1[i];
                    a2[u] = u2[i];
                    ia[u] = ui[i];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (i0 - u >= 0) {
                        const double zi = (yy[u] - a1[u] * z1 - a2[u] * z2) * ia[u];
                        out[u] = zi;
                        nrm += zi * zi;
                        z2 = z1;
                        z1 = zi;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (i0 - u >= 0) z[i0 - u] = out[u];
            }
            if (!(nrm > 0.0)) {
                for (int i = 0; i < m; ++i) z[i] = (i == (t + 1) % m) ? 1.0 : 0.0;
                nrm = 1.0;
            }
            scale = 1.0 / sqrt(nrm);
            if (extra > 0) --extra;
            if (extra < 0 && sqrt(nrm) >= grow) extra = 1;
        }
        for (int i = 0; i < m; ++i) z[i] *= scale;
    }
    __syncthreads();
    // modified Gram-Schmidt inside clusters (wave 0; entries lane and lane + 64)
    if (w == 0) {
        for (int i = 1; i < nwant; ++i) {
            const int cs = s.cstart[i];
            if (cs == i) continue;
            double* zi = s.z[i];
            const bool ia = lane < m, ib = lane + 64 < m;
            for (int j = cs; j < i; ++j) {
                const double* zj = s.z[j];
                double dp = 0.0;
                if (ia) dp += zj[lane] * zi[lane];
                if (ib) dp += zj[lane + 64] * zi[lane + 64];
                dp = wave_sum(dp);
                if (ia) zi[lane] -= dp * zj[lane];
                if (ib) zi[lane + 64] -= dp * zj[lane + 64];
                wave_lds_fence();
            }
            double nn = 0.0;
            if (ia) nn += zi[lane] * zi[lane];
            if (ib) nn += zi[lane + 64] * zi[lane + 64];
            nn = wave_sum(nn);
            const double in = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
            if (ia) zi[lane] *= in;
            if (ib) zi[lane + 64] *= in;
            wave_lds_fence();
        }
    }
    __syncthreads();

    // ---- 4. back-transformation ------------------------------------------------------
    const int vec = t >> 4, ls = t & 15;
    double2 x[kEM / 16];
#pragma unroll
    for (int q = 0; q < kEM / 16; ++q) {
        const int i = ls + 16 * q;
        x[q] = (vec < nwant && i < m) ? csc(s.ph[i], s.z[vec < nwant ? vec : 0][i]) : make_double2(0.0, 0.0);
    }
    for (int k1 = m - 2; k1 > 0; k1 -= kEStage) {  // reflectors k1 - 1 down to max(0, k1 - kEStage)
        const int k0 = max(0, k1 - kEStage);
        __syncthreads();  // the previous stage (or z) is no longer read
        for (int e = t; e < (k1 - k0) * kEM; e += kEThreads) {
            const int kk = e / kEM, i = e % kEM;
            const CT rv = refl[static_cast<size_t>(k0 + kk) * kEM + i];
            s.refl[kk][i] = make_double2(rv.x, rv.y);
        }
        __syncthreads();
        if (vec < nwant) {
            for (int k = k1 - 1; k >= k0; --k) {
                const double2* v = s.refl[k - k0];
                double2 sacc = make_double2(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < kEM / 16; ++q) cfma_conj(sacc, v[ls + 16 * q], x[q]);
                sacc.x = sum16_row(sacc.x);
                sacc.y = sum16_row(sacc.y);
                sacc = csc(sacc, 2.0);
#pragma unroll
                for (int q = 0; q < kEM / 16; ++q) x[q] = csub(x[q], cmul(v[ls + 16 * q], sacc));
            }
        }
    }
    if (vec < nwant) {
#pragma unroll
        for (int q = 0; q < kEM / 16; ++q) {
            const int i = ls + 16 * q;
            if (i < m) Y[i + static_cast<size_t>(vec) * ldy] = x[q];
        }
        if (ls == 0) theta[vec] = s.lam[vec];
    }
}

// T[i][j] for the new block columns j in [m0, m0 + nb) (and the mirrored rows) from
// P = V[:, 0:m0+nb]^H W[:, m0:m0+nb] (ld mn = m0 + nb); the new diagonal block is
// symmetrised, (P[m0+i, j] + conj(P[m0+j, i])) / 2, with a real diagonal.
__global__ void extend_projected_kernel(const double2* __restrict__ P, int m0, int nb, double2* __restrict__ T,
                                        int ldt) {
    const int mn = m0 + nb;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= mn) return;
    if (i < m0) {
        const double2 v = P[i + static_cast<size_t>(j) * mn];
        T[i + static_cast<size_t>(m0 + j) * ldt] = v;
        T[(m0 + j) + static_cast<size_t>(i) * ldt] = cconj(v);
        return;
    }
    const int ii = i - m0;
    if (ii > j) return;
    const double2 pij = P[(m0 + ii) + static_cast<size_t>(j) * mn];
    const double2 pji = P[(m0 + j) + static_cast<size_t>(ii) * mn];
    double2 avg = make_double2(0.5 * (pij.x + pji.x), 0.5 * (pij.y - pji.y));
    if (ii == j) avg.y = 0.0;
    T[(m0 + ii) + static_cast<size_t>(m0 + j) * ldt] = avg;
    T[(m0 + j) + static_cast<size_t>(m0 + ii) * ldt] = cconj(avg);
}

// T (ld x ld) = diag(theta[0..nb)), zero elsewhere.
__global__ void restart_projected_kernel(double2* __restrict__ T, int ld, const double* __restrict__ theta, int nb) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<int64_t>(ld) * ld) return;
    const int i = static_cast<int>(e % ld), j = static_cast<int>(e / ld);
    T[e] = (i == j && i < nb) ? make_double2(theta[i], 0.0) : make_double2(0.0, 0.0);
}

}  // namespace

int projected_eigen_max_dim() { return kEM; }
int projected_eigen_max_want() { return kEWant; }

void launch_projected_eigen(const double2* T, int ldt, int m, int nwant, double2* refl, double* theta, double2* Y,
                            int ldy, bool fp32_reduction, hipStream_t s) {
    if (m < 1 || m > kEM || nwant < 1 || nwant > kEWant || nwant > m || ldt < m || ldy < m)
        throw Error(XATU_ERR_INVALID_ARGUMENT, "projected eigensolver: unsupported dimensions");
    if (fp32_reduction) {
        hipLaunchKernelGGL(projected_eigen_kernel<float>, dim3(1), dim3(kEThreads), 0, s, T, ldt, m, nwant,
                           reinterpret_cast<float2*>(refl), theta, Y, ldy);
    } else {
        hipLaunchKernelGGL(projected_eigen_kernel<double>, dim3(1), dim3(kEThreads), 0, s, T, ldt, m, nwant, refl,
                           theta, Y, ldy);
    }
    XATU_HIP(hipGetLastError());
}

void launch_extend_projected(const double2* P, int m0, int nb, double2* T, int ldt, hipStream_t s) {
    const int mn = m0 + nb;
    if (nb <= 0 || mn > ldt) throw Error(XATU_ERR_INVALID_ARGUMENT, "projected matrix: block exceeds its size");
    hipLaunchKernelGGL(extend_projected_kernel, dim3((mn + 127) / 128, nb), dim3(128), 0, s, P, m0, nb, T, ldt);
    XATU_HIP(hipGetLastError());
}

void launch_restart_projected(double2* T, int ld, const double* theta, int nb, hipStream_t s) {
    const int64_t n = static_cast<int64_t>(ld) * ld;
    hipLaunchKernelGGL(restart_projected_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, T, ld,
                       theta, nb);
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
