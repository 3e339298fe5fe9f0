// Dense Hermitian eigensolver for small host-side problems (single-particle
// Bloch Hamiltonians, n <= ~64, and Davidson Rayleigh-Ritz matrices, n <= ~300).
//
// Method: Householder reduction of the Hermitian matrix to a Hermitian
// tridiagonal T = Q^H A Q, a diagonal unitary phase change that makes T real
// symmetric, implicit QL iterations with Wilkinson-type shifts on the real
// tridiagonal (accumulating the real rotations), then back-transformation of
// the eigenvectors through the phases and the reflectors. Eigenvalues are
// returned in ascending order, as LAPACK zheevd returns them (the reference
// calls zheevd through arma::eig_sym, SystemTB.cpp:124, ExcitonTB.cpp:942).
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace xatu_linalg {

using cd = std::complex<double>;

// Eigen-decomposition of the Hermitian n x n matrix A (column-major, leading
// dimension lda; only the lower triangle and the real part of the diagonal are
// read). On return w[0..n) holds the eigenvalues in ascending order and, if Z
// is non-null, Z (column-major, leading dimension n) holds the orthonormal
// eigenvectors as columns.
inline void heev(int n, const cd* Ain, int lda, double* w, cd* Z) {
    if (n <= 0) return;
    std::vector<cd> A(static_cast<size_t>(n) * n);
    for (int j = 0; j < n; ++j) {
        for (int i = j; i < n; ++i) {
            cd v = Ain[i + static_cast<size_t>(j) * lda];
            if (i == j) v = cd(v.real(), 0.0);
            A[i + static_cast<size_t>(j) * n] = v;
            A[j + static_cast<size_t>(i) * n] = std::conj(v);
        }
    }
    auto at = [&](int i, int j) -> cd& { return A[i + static_cast<size_t>(j) * n]; };

    std::vector<double> d(n, 0.0), e(n, 0.0);
    std::vector<cd> ecplx(n, cd(0.0, 0.0));
    // Householder vectors: step k stores v_k (length n-k-1) at offset voff[k].
    std::vector<size_t> voff(n, 0);
    std::vector<cd> vstore;
    std::vector<char> vused(n, 0);
    std::vector<cd> p(n), wv(n);

    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;
        double tail2 = 0.0;
        for (int i = 1; i < m; ++i) tail2 += std::norm(at(k + 1 + i, k));
        double xnorm2 = tail2 + std::norm(at(k + 1, k));
        voff[k] = vstore.size();
        vstore.resize(vstore.size() + m);
        cd* v = vstore.data() + voff[k];
        if (xnorm2 == 0.0 || (tail2 <= 0.0)) {
            // Column already tridiagonal below the sub-diagonal: no reflection.
            ecplx[k] = at(k + 1, k);
            d[k] = at(k, k).real();
            continue;
        }
        const double xnorm = std::sqrt(xnorm2);
        cd x0 = at(k + 1, k);
        double ax0 = std::abs(x0);
        cd phase = (ax0 > 0.0) ? x0 / ax0 : cd(1.0, 0.0);
        cd alpha = -phase * xnorm;
        for (int i = 0; i < m; ++i) v[i] = at(k + 1 + i, k);
        v[0] -= alpha;
        double vn2 = 0.0;
        for (int i = 0; i < m; ++i) vn2 += std::norm(v[i]);
        double vn = std::sqrt(vn2);
        for (int i = 0; i < m; ++i) v[i] /= vn;
        vused[k] = 1;

        // Trailing block B = A[k+1:, k+1:]; B <- (I - 2vv^H) B (I - 2vv^H)
        // = B - 2 (v w^H + w v^H), with p = B v, c = v^H p, w = p - c v.
        for (int i = 0; i < m; ++i) {
            cd s(0.0, 0.0);
            for (int j = 0; j < m; ++j) s += at(k + 1 + i, k + 1 + j) * v[j];
            p[i] = s;
        }
        cd c(0.0, 0.0);
        for (int i = 0; i < m; ++i) c += std::conj(v[i]) * p[i];
        double cr = c.real();
        for (int i = 0; i < m; ++i) wv[i] = p[i] - cr * v[i];
        for (int j = 0; j < m; ++j) {
            for (int i = 0; i < m; ++i) {
                at(k + 1 + i, k + 1 + j) -= 2.0 * (v[i] * std::conj(wv[j]) + wv[i] * std::conj(v[j]));
            }
        }
        at(k + 1, k) = alpha;
        at(k, k + 1) = std::conj(alpha);
        for (int i = 1; i < m; ++i) {
            at(k + 1 + i, k) = cd(0.0, 0.0);
            at(k, k + 1 + i) = cd(0.0, 0.0);
        }
        ecplx[k] = alpha;
        d[k] = at(k, k).real();
    }
    if (n >= 2) {
        d[n - 2] = at(n - 2, n - 2).real();
        ecplx[n - 2] = at(n - 1, n - 2);
    }
    d[n - 1] = at(n - 1, n - 1).real();

    // Phases that make the tridiagonal real: ph_{k+1} = ph_k * e_k / |e_k|.
    std::vector<cd> ph(n, cd(1.0, 0.0));
    for (int k = 0; k + 1 < n; ++k) {
        double ae = std::abs(ecplx[k]);
        e[k] = ae;
        ph[k + 1] = (ae > 0.0) ? ph[k] * (ecplx[k] / ae) : ph[k];
    }
    e[n - 1] = 0.0;

    // Implicit QL on the real symmetric tridiagonal (d, e), accumulating
    // rotations into the real matrix zr (column i = eigenvector i).
    const bool want = (Z != nullptr);
    std::vector<double> zr;
    if (want) {
        zr.assign(static_cast<size_t>(n) * n, 0.0);
        for (int i = 0; i < n; ++i) zr[i + static_cast<size_t>(i) * n] = 1.0;
    }
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m != l) {
                if (iter++ == 200) throw std::runtime_error("heev: QL iteration did not converge");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
                double s = 1.0, c = 1.0, pp = 0.0;
                int i;
                bool deflated = false;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    double b = c * e[i];
                    r = std::hypot(f, g);
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
                    if (want) {
                        double* zi = zr.data() + static_cast<size_t>(i) * n;
                        double* zi1 = zr.data() + static_cast<size_t>(i + 1) * n;
                        for (int kk = 0; kk < n; ++kk) {
                            double t = zi1[kk];
                            zi1[kk] = s * zi[kk] + c * t;
                            zi[kk] = c * zi[kk] - s * t;
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

    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int i = 0; i < n; ++i) w[i] = d[order[i]];
    if (!want) return;

    // Eigenvectors: x = H_0 H_1 ... H_{n-3} Ph z.
    for (int col = 0; col < n; ++col) {
        const double* zc = zr.data() + static_cast<size_t>(order[col]) * n;
        cd* x = Z + static_cast<size_t>(col) * n;
        for (int i = 0; i < n; ++i) x[i] = ph[i] * zc[i];
        for (int k = n - 3; k >= 0; --k) {
            if (!vused[k]) continue;
            const int m = n - k - 1;
            const cd* v = vstore.data() + voff[k];
            cd s(0.0, 0.0);
            for (int i = 0; i < m; ++i) s += std::conj(v[i]) * x[k + 1 + i];
            s *= 2.0;
            for (int i = 0; i < m; ++i) x[k + 1 + i] -= s * v[i];
        }
    }
}

}  // namespace xatu_linalg
