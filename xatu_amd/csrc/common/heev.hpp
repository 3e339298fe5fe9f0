// Dense Hermitian eigensolver for small host-side problems (single-particle
// Bloch Hamiltonians, n <= ~64, and Davidson Rayleigh-Ritz matrices, n <= ~300).
//
// Method: Householder reduction of the Hermitian matrix to a Hermitian
// tridiagonal T = Q^H A Q, a diagonal unitary phase change that makes T real
// symmetric, then either
//  - heev:        implicit QL iterations with the rotations accumulated (all
//                 eigenvectors), or
//  - heev_lowest: QL eigenvalues only, and the wanted lowest eigenvectors of the
//                 real tridiagonal by inverse iteration (LU with partial pivoting,
//                 vectors re-orthogonalised inside clusters of close eigenvalues),
// and the back-transformation through the phases and the reflectors. Eigenvalues
// are returned in ascending order, as LAPACK zheevd returns them (the reference
// calls zheevd through arma::eig_sym, SystemTB.cpp:124, ExcitonTB.cpp:942).
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <numeric>
#include <stdexcept>
#include <vector>

// The O(n^3) loops below are also compiled for AVX2/FMA (x86-64 hosts of this GPU have
// it); the loader picks the clone the host CPU supports.
#if defined(__x86_64__) && !defined(__HIP_DEVICE_COMPILE__)
#define XATU_HOST_CLONES __attribute__((target_clones("arch=haswell", "default")))
#else
#define XATU_HOST_CLONES
#endif
// The dot products of these loops may be summed in any order (vectorised partial sums).
#if defined(__clang__)
#define XATU_REASSOC _Pragma("clang fp reassociate(on)")
#else
#define XATU_REASSOC
#endif

namespace xatu_linalg {

using cd = std::complex<double>;

namespace detail {

// Result of the Householder stage: real tridiagonal (d, e), phases and reflectors.
struct Tridiag {
    int n = 0;
    std::vector<double> d, e;     // e[k] couples k and k+1; e[n-1] = 0
    std::vector<cd> ph;           // T_complex = Ph T_real Ph^H
    std::vector<double> vre, vim; // reflector k stored at offset voff[k], length n-k-1
    std::vector<size_t> voff;
    std::vector<char> vused;
};

// sqrt(a^2 + b^2) without the cost of std::hypot where no scaling is needed.
inline double fast_hypot(double a, double b) {
    a = std::fabs(a);
    b = std::fabs(b);
    const double m = std::max(a, b);
    if (m > 1e150 || (m < 1e-150 && m > 0.0)) return std::hypot(a, b);
    return std::sqrt(a * a + b * b);
}

// Two-sided reflection of the trailing Hermitian block B (m x m, lower triangle stored
// split in br/bi with leading dimension ld) by I - 2 v v^H: p = B v over the lower
// triangle (the stored B_ij, i > j, contributes B_ij v_j to p_i and conj(B_ij) v_i to
// p_j), w = p - (v^H p) v, then the lower triangle -= 2 (v w^H + w v^H).
XATU_HOST_CLONES static void reflect_trailing(int m, int ld, double* __restrict__ ar, double* __restrict__ ai,
                                              const double* __restrict__ vr, const double* __restrict__ vi,
                                              double* __restrict__ pr, double* __restrict__ pi,
                                              double* __restrict__ wr, double* __restrict__ wi) {
    XATU_REASSOC
    for (int i = 0; i < m; ++i) {
        pr[i] = 0.0;
        pi[i] = 0.0;
    }
    for (int j = 0; j < m; ++j) {
        const double* br = ar + static_cast<size_t>(j) * ld;
        const double* bi = ai + static_cast<size_t>(j) * ld;
        const double vjr = vr[j], vji = vi[j];
        double sr = br[j] * vjr, si = br[j] * vji;  // real diagonal
        for (int i = j + 1; i < m; ++i) {
            pr[i] += br[i] * vjr - bi[i] * vji;
            pi[i] += br[i] * vji + bi[i] * vjr;
            sr += br[i] * vr[i] + bi[i] * vi[i];
            si += br[i] * vi[i] - bi[i] * vr[i];
        }
        pr[j] += sr;
        pi[j] += si;
    }
    double cr = 0.0;  // c = v^H p (real)
    for (int i = 0; i < m; ++i) cr += vr[i] * pr[i] + vi[i] * pi[i];
    for (int i = 0; i < m; ++i) {
        wr[i] = pr[i] - cr * vr[i];
        wi[i] = pi[i] - cr * vi[i];
    }
    for (int j = 0; j < m; ++j) {
        double* br = ar + static_cast<size_t>(j) * ld;
        double* bi = ai + static_cast<size_t>(j) * ld;
        const double wjr = 2.0 * wr[j], wji = -2.0 * wi[j];  // 2 conj(w_j)
        const double vjr = 2.0 * vr[j], vji = -2.0 * vi[j];  // 2 conj(v_j)
        for (int i = j; i < m; ++i) {
            br[i] -= vr[i] * wjr - vi[i] * wji + wr[i] * vjr - wi[i] * vji;
            bi[i] -= vr[i] * wji + vi[i] * wjr + wr[i] * vji + wi[i] * vjr;
        }
        bi[j] = 0.0;  // the diagonal stays real
    }
}

// x (n x nz block, split, leading dimension ld) -= 2 v (v^H x) for a unit reflector v of
// length m applied to rows [0, m) of x.
XATU_HOST_CLONES static void reflect_block(int m, int nz, int ld, const double* __restrict__ vr,
                                           const double* __restrict__ vi, double* __restrict__ xr,
                                           double* __restrict__ xi) {
    XATU_REASSOC
    for (int c = 0; c < nz; ++c) {
        double* ar = xr + static_cast<size_t>(c) * ld;
        double* ai = xi + static_cast<size_t>(c) * ld;
        double sr = 0.0, si = 0.0;  // s = v^H x
        for (int i = 0; i < m; ++i) {
            sr += vr[i] * ar[i] + vi[i] * ai[i];
            si += vr[i] * ai[i] - vi[i] * ar[i];
        }
        sr *= 2.0;
        si *= 2.0;
        for (int i = 0; i < m; ++i) {
            ar[i] -= sr * vr[i] - si * vi[i];
            ai[i] -= sr * vi[i] + si * vr[i];
        }
    }
}

// Householder reduction on the lower triangle only (as LAPACK zhetd2 with uplo = L):
// step k reflects column k below the diagonal, forms p = B v with B the trailing
// Hermitian block (one pass over its lower triangle) and applies the rank-2 update
// B -= 2 (v w^H + w v^H), w = p - (v^H p) v, to the lower triangle.
inline void tridiagonalize(int n, const cd* Ain, int lda, Tridiag& t) {
    t.n = n;
    // split storage of the lower triangle, column-major
    std::vector<double> ar(static_cast<size_t>(n) * n), ai(static_cast<size_t>(n) * n);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) {
            const cd v = Ain[i + static_cast<size_t>(j) * lda];
            ar[i + static_cast<size_t>(j) * n] = v.real();
            ai[i + static_cast<size_t>(j) * n] = (i == j) ? 0.0 : v.imag();
        }
    t.d.assign(n, 0.0);
    t.e.assign(n, 0.0);
    std::vector<cd> ec(n, cd(0.0, 0.0));
    t.voff.assign(n, 0);
    t.vused.assign(n, 0);
    t.vre.clear();
    t.vim.clear();
    t.vre.reserve(static_cast<size_t>(n) * n / 2 + n);
    t.vim.reserve(static_cast<size_t>(n) * n / 2 + n);
    std::vector<double> pr(n), pi(n), wr(n), wi(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;
        const size_t col = static_cast<size_t>(k) * n + k + 1;  // start of x = A[k+1:, k]
        double tail2 = 0.0;
        for (int i = 1; i < m; ++i) tail2 += ar[col + i] * ar[col + i] + ai[col + i] * ai[col + i];
        const double x0r = ar[col], x0i = ai[col];
        const double xnorm2 = tail2 + x0r * x0r + x0i * x0i;
        t.voff[k] = t.vre.size();
        t.vre.resize(t.vre.size() + m);
        t.vim.resize(t.vim.size() + m);
        if (xnorm2 == 0.0 || tail2 <= 0.0) {
            ec[k] = cd(x0r, x0i);
            t.d[k] = ar[static_cast<size_t>(k) * n + k];
            continue;
        }
        const double xnorm = std::sqrt(xnorm2);
        const double ax0 = std::hypot(x0r, x0i);
        const double phr = ax0 > 0.0 ? x0r / ax0 : 1.0, phi = ax0 > 0.0 ? x0i / ax0 : 0.0;
        const double alr = -phr * xnorm, ali = -phi * xnorm;  // alpha = -phase |x|
        double* vr = t.vre.data() + t.voff[k];
        double* vi = t.vim.data() + t.voff[k];
        for (int i = 0; i < m; ++i) {
            vr[i] = ar[col + i];
            vi[i] = ai[col + i];
        }
        vr[0] -= alr;
        vi[0] -= ali;
        double vn2 = 0.0;
        for (int i = 0; i < m; ++i) vn2 += vr[i] * vr[i] + vi[i] * vi[i];
        const double inv = 1.0 / std::sqrt(vn2);
        for (int i = 0; i < m; ++i) {
            vr[i] *= inv;
            vi[i] *= inv;
        }
        t.vused[k] = 1;
        const size_t base = static_cast<size_t>(k + 1) * n + k + 1;
        reflect_trailing(m, n, ar.data() + base, ai.data() + base, vr, vi, pr.data(), pi.data(), wr.data(), wi.data());
        ec[k] = cd(alr, ali);
        t.d[k] = ar[static_cast<size_t>(k) * n + k];
    }
    if (n >= 2) {
        t.d[n - 2] = ar[static_cast<size_t>(n - 2) * n + n - 2];
        ec[n - 2] = cd(ar[static_cast<size_t>(n - 2) * n + n - 1], ai[static_cast<size_t>(n - 2) * n + n - 1]);
    }
    t.d[n - 1] = ar[static_cast<size_t>(n - 1) * n + n - 1];
    t.ph.assign(n, cd(1.0, 0.0));
    for (int k = 0; k + 1 < n; ++k) {
        const double ae = std::abs(ec[k]);
        t.e[k] = ae;
        t.ph[k + 1] = (ae > 0.0) ? t.ph[k] * (ec[k] / ae) : t.ph[k];
    }
    t.e[n - 1] = 0.0;
}

// Implicit QL on (d, e); if zr is non-null the rotations are accumulated into it
// (n x n, column i = eigenvector i). d receives the (unsorted) eigenvalues.
inline void tql(int n, std::vector<double>& d, std::vector<double> e, double* zr) {
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m != l) {
                if (iter++ == 200) throw std::runtime_error("heev: QL iteration did not converge");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = fast_hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, pp = 0.0;
                bool deflated = false;
                for (int i = m - 1; i >= l; --i) {
                    const double f = s * e[i];
                    const double b = c * e[i];
                    r = fast_hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= pp;
                        e[m] = 0.0;
                        deflated = true;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - pp;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    pp = s * r;
                    d[i + 1] = g + pp;
                    g = c * r - b;
                    if (zr) {
                        double* zi = zr + static_cast<size_t>(i) * n;
                        double* zi1 = zr + static_cast<size_t>(i + 1) * n;
                        for (int kk = 0; kk < n; ++kk) {
                            const double tt = zi1[kk];
                            zi1[kk] = s * zi[kk] + c * tt;
                            zi[kk] = c * zi[kk] - s * tt;
                        }
                    }
                }
                if (deflated) continue;
                d[l] -= pp;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
}

// x (length n) <- H_0 H_1 ... H_{n-3} Ph z  for a real tridiagonal eigenvector z.
inline void back_transform(const Tridiag& t, const double* z, cd* x) {
    const int n = t.n;
    for (int i = 0; i < n; ++i) x[i] = t.ph[i] * z[i];
    for (int k = n - 3; k >= 0; --k) {
        if (!t.vused[k]) continue;
        const int m = n - k - 1;
        const double* vr = t.vre.data() + t.voff[k];
        const double* vi = t.vim.data() + t.voff[k];
        double sr = 0.0, si = 0.0;  // s = v^H x
        for (int i = 0; i < m; ++i) {
            const cd xi = x[k + 1 + i];
            sr += vr[i] * xi.real() + vi[i] * xi.imag();
            si += vr[i] * xi.imag() - vi[i] * xi.real();
        }
        sr *= 2.0;
        si *= 2.0;
        for (int i = 0; i < m; ++i) {
            x[k + 1 + i] -= cd(sr * vr[i] - si * vi[i], sr * vi[i] + si * vr[i]);
        }
    }
}

// The same for nz vectors at once (z: n x nz, x: n x nz, column-major): each reflector is
// applied to all columns while it is in cache.
inline void back_transform_block(const Tridiag& t, const double* z, int nz, cd* x) {
    const int n = t.n;
    std::vector<double> xr(static_cast<size_t>(n) * nz), xi(static_cast<size_t>(n) * nz);
    for (int c = 0; c < nz; ++c)
        for (int i = 0; i < n; ++i) {
            const cd v = t.ph[i] * z[i + static_cast<size_t>(c) * n];
            xr[i + static_cast<size_t>(c) * n] = v.real();
            xi[i + static_cast<size_t>(c) * n] = v.imag();
        }
    for (int k = n - 3; k >= 0; --k) {
        if (!t.vused[k]) continue;
        const int m = n - k - 1;
        reflect_block(m, nz, n, t.vre.data() + t.voff[k], t.vim.data() + t.voff[k], xr.data() + k + 1,
                      xi.data() + k + 1);
    }
    for (int c = 0; c < nz; ++c)
        for (int i = 0; i < n; ++i)
            x[i + static_cast<size_t>(c) * n] = cd(xr[i + static_cast<size_t>(c) * n], xi[i + static_cast<size_t>(c) * n]);
}

// Eigenvector of the real symmetric tridiagonal (d, e) for eigenvalue lam by inverse
// iteration; z is orthogonalised against the columns of prev (a cluster).
inline void inverse_iteration(const std::vector<double>& d, const std::vector<double>& e, double lam, double tnorm,
                              const std::vector<const double*>& prev, unsigned seed, double* z) {
    const int n = static_cast<int>(d.size());
    const double eps = DBL_EPSILON * std::max(tnorm, DBL_MIN);
    // LU with partial pivoting of T - lam I: rows of U hold (u0, u1, u2) on the diagonal
    // and two super-diagonals; l[i] multiplier; piv[i] whether rows i, i+1 were swapped.
    std::vector<double> u0(n), u1(n, 0.0), u2(n, 0.0), lm(n, 0.0);
    std::vector<char> piv(n, 0);
    auto guard = [&](double x) { return (std::fabs(x) < eps) ? (x < 0.0 ? -eps : eps) : x; };
    {
        // pending row: alpha at column i, beta at column i+1
        double alpha = d[0] - lam, beta = (n > 1) ? e[0] : 0.0;
        for (int i = 0; i < n - 1; ++i) {
            const double sub = e[i];           // T[i+1, i]
            const double nd = d[i + 1] - lam;  // T[i+1, i+1]
            const double ns = (i + 2 < n) ? e[i + 1] : 0.0;
            if (std::fabs(alpha) >= std::fabs(sub)) {
                alpha = guard(alpha);
                const double mult = sub / alpha;
                lm[i] = mult;
                u0[i] = alpha;
                u1[i] = beta;
                u2[i] = 0.0;
                alpha = nd - mult * beta;
                beta = ns;
            } else {
                const double mult = alpha / sub;
                lm[i] = mult;
                piv[i] = 1;
                u0[i] = sub;
                u1[i] = nd;
                u2[i] = ns;
                alpha = beta - mult * nd;
                beta = -mult * ns;
            }
        }
        u0[n - 1] = guard(alpha);
    }
    // start vector: fixed pseudo-random
    unsigned s = seed * 2654435761u + 12345u;
    for (int i = 0; i < n; ++i) {
        s = s * 1664525u + 1013904223u;
        z[i] = 0.5 + (s >> 8) * (1.0 / 16777216.0);
    }
    std::vector<double> y(n);
    for (int it = 0; it < 4; ++it) {
        // solve L y = z (apply the row operations), then U x = y
        for (int i = 0; i < n; ++i) y[i] = z[i];
        for (int i = 0; i < n - 1; ++i) {
            if (piv[i]) std::swap(y[i], y[i + 1]);
            y[i + 1] -= lm[i] * y[i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double v = y[i];
            if (i + 1 < n) v -= u1[i] * z[i + 1];
            if (i + 2 < n) v -= u2[i] * z[i + 2];
            z[i] = v / u0[i];
        }
        for (const double* p : prev) {
            double dot = 0.0;
            for (int i = 0; i < n; ++i) dot += p[i] * z[i];
            for (int i = 0; i < n; ++i) z[i] -= dot * p[i];
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm += z[i] * z[i];
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0)) {
            for (int i = 0; i < n; ++i) z[i] = (i == static_cast<int>(seed % n)) ? 1.0 : 0.0;
            nrm = 1.0;
        }
        for (int i = 0; i < n; ++i) z[i] /= nrm;
    }
}

}  // namespace detail

// Eigen-decomposition of the Hermitian n x n matrix A (column-major, leading
// dimension lda; only the lower triangle and the real part of the diagonal are
// read). On return w[0..n) holds the eigenvalues in ascending order and, if Z
// is non-null, Z (column-major, leading dimension n) holds the orthonormal
// eigenvectors as columns.
inline void heev(int n, const cd* Ain, int lda, double* w, cd* Z) {
    if (n <= 0) return;
    detail::Tridiag t;
    detail::tridiagonalize(n, Ain, lda, t);
    std::vector<double> d = t.d;
    std::vector<double> zr;
    if (Z) {
        zr.assign(static_cast<size_t>(n) * n, 0.0);
        for (int i = 0; i < n; ++i) zr[i + static_cast<size_t>(i) * n] = 1.0;
    }
    detail::tql(n, d, t.e, Z ? zr.data() : nullptr);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int i = 0; i < n; ++i) w[i] = d[order[i]];
    if (!Z) return;
    for (int c = 0; c < n; ++c) detail::back_transform(t, zr.data() + static_cast<size_t>(order[c]) * n, Z + static_cast<size_t>(c) * n);
}

// All eigenvalues (ascending, w[0..n)) and the eigenvectors of the nwant lowest
// (Z: n x nwant column-major).
inline void heev_lowest(int n, const cd* Ain, int lda, int nwant, double* w, cd* Z) {
    if (n <= 0) return;
    nwant = std::min(nwant, n);
    detail::Tridiag t;
    detail::tridiagonalize(n, Ain, lda, t);
    std::vector<double> d = t.d;
    detail::tql(n, d, t.e, nullptr);
    std::sort(d.begin(), d.end());
    for (int i = 0; i < n; ++i) w[i] = d[i];
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i) tnorm = std::max(tnorm, std::fabs(t.d[i]) + std::fabs(t.e[i]) + (i ? std::fabs(t.e[i - 1]) : 0.0));
    const double cluster = 1e-3 * std::max(tnorm, 1e-300);
    std::vector<double> zs(static_cast<size_t>(n) * nwant);
    int cstart = 0;
    for (int i = 0; i < nwant; ++i) {
        if (i > 0 && w[i] - w[i - 1] > cluster) cstart = i;
        std::vector<const double*> prev;
        for (int j = cstart; j < i; ++j) prev.push_back(zs.data() + static_cast<size_t>(j) * n);
        // separate coincident shifts inside a cluster slightly, as LAPACK stein does
        double lam = w[i];
        if (i > cstart && lam - w[i - 1] < 10.0 * DBL_EPSILON * tnorm) lam = w[i - 1] + 10.0 * DBL_EPSILON * tnorm;
        detail::inverse_iteration(t.d, t.e, lam, tnorm, prev, static_cast<unsigned>(i + 1),
                                  zs.data() + static_cast<size_t>(i) * n);
    }
    detail::back_transform_block(t, zs.data(), nwant, Z);
}

}  // namespace xatu_linalg
