// Device evaluation of the screened interactions in FP64.
//   Keldysh: V(r) = e / (8e-10 eps0 eps_bar r0) * (H0(r/r0) - Y0(r/r0))   [eV, r in Angstrom]
//   Coulomb: V(r) = e / (4e-10 pi eps0 r)
// with the reference's special cases: r == 0 evaluates at the regularization
// distance, r > cutoff gives 0 (src/ExcitonTB.cpp:347-381), and its constants
// (include/xatu/ExcitonTB.hpp:10-14).
#pragma once

#include <hip/hip_runtime.h>

#include "context.hpp"

namespace xatu_dev {

constexpr double kPi = 3.141592653589793;
constexpr double kEc = 1.6021766E-19;
constexpr double kEps0 = 8.8541878E-12;

// Struve function H0(x), x >= 0, following the published Zhang & Jin (1996)
// STVH0 scheme the reference uses (src/ExcitonTB.cpp:306-337):
//  x <= 20: power series H0 = (2x/pi) sum_k r_k, r_k = -r_{k-1} x^2/(2k+1)^2,
//           at most 60 terms, stopping once |r_k| < 1e-12 |sum|;
//  x  > 20: H0 = 2/(pi x) sum_k s_k + Y0(x), with the asymptotic terms
//           s_k = -s_{k-1} ((2k-1)/x)^2 (at most int((x+1)/2) terms, 25 for x >= 50)
//           and the rational approximation of Y0 with the published coefficients.
__host__ __device__ inline double struve_h0(double x) {
    double s = 1.0, r = 1.0;
    if (x <= 20.0) {
        const double a0 = 2.0 * x / kPi;
        for (int k = 1; k < 61; ++k) {
            const double d = 2.0 * k + 1.0;
            r = -r * x / d * x / d;
            s += r;
            if (fabs(r) < fabs(s) * 1.0e-12) break;
        }
        return a0 * s;
    }
    int km = static_cast<int>(0.5 * (x + 1.0));
    if (x >= 50.0) km = 25;
    for (int k = 1; k <= km; ++k) {
        const double t = (2.0 * k - 1.0) / x;
        r = -r * (t * t);
        s += r;
        if (fabs(r) < fabs(s) * 1.0e-12) break;
    }
    const double t = 4.0 / x;
    const double t2 = t * t;
    const double p0 =
        ((((-.37043e-5 * t2 + .173565e-4) * t2 - .487613e-4) * t2 + .17343e-3) * t2 - 0.1753062e-2) * t2 +
        .3989422793;
    const double q0 =
        t * (((((.32312e-5 * t2 - 0.142078e-4) * t2 + 0.342468e-4) * t2 - 0.869791e-4) * t2 + 0.4564324e-3) *
                 t2 -
             0.0124669441);
    const double ta0 = x - 0.25 * kPi;
    const double by0 = 2.0 / sqrt(x) * (p0 * sin(ta0) + q0 * cos(ta0));
    return 2.0 / (kPi * x) * s + by0;
}

// |R + tau_a - tau_b| for the cell R = sum_j m_j a_j (integer coefficients m, padded to 3),
// rounded operation by operation, without fused multiply-adds, in a fixed order:
// R = m_0 a_1 + m_1 a_2 (+ m_2 a_3), d = (R + tau_a) - tau_b, r = sqrt((dx^2 + dy^2) + dz^2).
// The Keldysh potential switches between two approximations of H0 at r / r0 = 20 (they
// differ by ~1e-8 there), so a lattice vector that lies exactly on that radius must round
// the same way wherever its distance is evaluated.
__device__ inline double cell_distance(const LatticeDesc& lat, const int* m, int a, int b) {
#pragma clang fp contract(off)
    double v[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < lat.ndim; ++j) {
        const double mj = static_cast<double>(m[j]);
        for (int x = 0; x < 3; ++x) v[x] = v[x] + mj * lat.bravais[3 * j + x];
    }
    double d[3];
    for (int x = 0; x < 3; ++x) d[x] = (v[x] + lat.tau[3 * a + x]) - lat.tau[3 * b + x];
    return sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
}

__device__ inline double keldysh_pref(const PotDesc& p) {
    return kEc / (8E-10 * kEps0 * p.eps_bar * p.r0);
}

// V(r) in eV for the configured potential.
__device__ inline double potential(double r, const PotDesc& p) {
    if (p.kind == XATU_POTENTIAL_COULOMB) {
        const double ar = fabs(r);
        if (ar > p.cutoff) return 0.0;
        return (ar != 0.0) ? kEc / (4E-10 * kPi * kEps0 * ar) : kEc * 1E10 / (4 * kPi * kEps0 * p.reg);
    }
    if (r == 0.0) {
        const double x = p.reg / p.r0;
        return keldysh_pref(p) * (struve_h0(x) - y0(x));
    }
    if (r > p.cutoff) return 0.0;
    const double x = fabs(r) / p.r0;
    return keldysh_pref(p) * (struve_h0(x) - y0(x));
}

// Struve H1(x) by its power series, H1 = (2 x^2 / (3 pi)) sum_k u_k with
// u_k = -u_{k-1} x^2 / ((2k+1)(2k+3)); the terms grow to ~e^x before they cancel, so it
// is used for x <= 5 only (relative error below 1e-14 there).
__host__ __device__ inline double struve_h1_series(double x) {
    double s = 1.0, u = 1.0;
    for (int k = 1; k < 80; ++k) {
        u = -u * x * x / ((2.0 * k + 1.0) * (2.0 * k + 3.0));
        s += u;
        if (fabs(u) < fabs(s) * 1.0e-16) break;
    }
    return 2.0 * x * x / (3.0 * kPi) * s;
}

// Gauss-Legendre rule of kGlNodes points on [0, kGlLength]: nodes then weights.
constexpr int kGlNodes = 48;
constexpr double kGlLength = 40.0;

// Nodes and weights of the Gauss-Legendre rule above, by Newton's iteration on the
// Legendre polynomial (three-term recurrence) from the usual cosine estimates of its roots.
inline void gauss_legendre_rule(double* out /* 2 * kGlNodes */) {
    const int n = kGlNodes;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(kPi * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        const double w = 2.0 / ((1.0 - z * z) * pp * pp);
        // map [-1, 1] to [0, L]
        out[i] = 0.5 * kGlLength * (1.0 - z);
        out[n - 1 - i] = 0.5 * kGlLength * (1.0 + z);
        out[n + i] = out[2 * n - 1 - i] = 0.5 * kGlLength * w;
    }
}

// g(x) = d(x f)/dx = f + x f' for f = STVH0(x) - Y0(x), the function V evaluates.
//  x <= 5:      f + x (2/pi - H1 + Y1) with the H1 power series;
//  5 < x <= 20: f and x f' cancel to O(x^-3) while each is O(1/x), so g comes from the
//               integral form f(x) = (2/pi) int_0^inf e^{-xt} / sqrt(1 + t^2) dt, which gives
//               g(x) = (2/pi) int_0^inf e^{-u} (u^2 / x^3) (1 + u^2/x^2)^(-3/2) du, a positive
//               integrand, on the Gauss-Legendre rule (gl: nodes, then weights) over [0, 40]
//               (the tail beyond is below e^-40 relative);
//  x > 20:      the exact derivative of the asymptotic branch of struve_h0 minus Y0: with
//               H0 = 2/(pi x) S + BY0 the S/x terms cancel analytically, leaving
//               g = (2/pi) S'(x) + (BY0 - Y0) + x (BY0' + Y1) over the same series terms, so
//               g is the derivative of the potential actually evaluated there.
// Relative error against 40-digit values for x <= 20: below 4e-13 (series), 1e-14 (integral).
__host__ __device__ inline double keldysh_dxf(double x, const double* gl) {
    if (x <= 5.0) {
        return struve_h0(x) - y0(x) + x * (2.0 / kPi - struve_h1_series(x) + y1(x));
    }
    if (x <= 20.0) {
        const double ix = 1.0 / x;
        double s = 0.0;
        for (int i = 0; i < kGlNodes; ++i) {
            const double u = gl[i];
            const double q = 1.0 + (u * ix) * (u * ix);
            s += gl[kGlNodes + i] * exp(-u) * (u * u) / (q * sqrt(q));
        }
        return 2.0 / kPi * s * ix * ix * ix;
    }
    int km = static_cast<int>(0.5 * (x + 1.0));
    if (x >= 50.0) km = 25;
    double S = 1.0, r = 1.0, dS = 0.0;
    for (int k = 1; k <= km; ++k) {
        const double tk = (2.0 * k - 1.0) / x;
        r = -r * (tk * tk);  // (-1)^k ((2k-1)!!)^2 x^-2k
        S += r;
        dS += -2.0 * k * r / x;
        if (fabs(r) < fabs(S) * 1.0e-12) break;
    }
    // BY0 = 2 x^-1/2 (P(t^2) sin th + t Qp(t^2) cos th), t = 4/x, th = x - pi/4
    const double t = 4.0 / x, t2 = t * t, dt = -t * t / 4.0;
    const double P =
        ((((-.37043e-5 * t2 + .173565e-4) * t2 - .487613e-4) * t2 + .17343e-3) * t2 - 0.1753062e-2) * t2 +
        .3989422793;
    const double dP = (((5 * -.37043e-5 * t2 + 4 * .173565e-4) * t2 - 3 * .487613e-4) * t2 + 2 * .17343e-3) * t2 -
                      0.1753062e-2;  // dP / d(t^2)
    const double Qp = ((((.32312e-5 * t2 - 0.142078e-4) * t2 + 0.342468e-4) * t2 - 0.869791e-4) * t2 +
                       0.4564324e-3) * t2 -
                      0.0124669441;
    const double dQp = (((5 * .32312e-5 * t2 - 4 * 0.142078e-4) * t2 + 3 * 0.342468e-4) * t2 - 2 * 0.869791e-4) * t2 +
                       0.4564324e-3;  // dQp / d(t^2)
    const double Q = t * Qp;
    const double dPx = dP * 2.0 * t * dt;
    const double dQx = (Qp + t * dQp * 2.0 * t) * dt;
    const double th = x - 0.25 * kPi, sn = sin(th), cs = cos(th);
    const double rx = 1.0 / sqrt(x);
    const double by0 = 2.0 * rx * (P * sn + Q * cs);
    const double dby0 = -rx / x * (P * sn + Q * cs) + 2.0 * rx * (dPx * sn + P * cs + dQx * cs - Q * sn);
    return 2.0 / kPi * dS + (by0 - y0(x)) + x * (dby0 + y1(x));
}

// dV/dr0 for the Keldysh potential at distance r (same special cases as V):
// with f = H0 - Y0 and x = r/r0, V = C f(x) / (eps_bar r0), so
// dV/dr0 = -C / (eps_bar r0^2) * (f(x) + x f'(x)) = -C / (eps_bar r0^2) * g(x).
__device__ inline double keldysh_dr0(double r, const PotDesc& p, const double* gl) {
    double x;
    if (r == 0.0) {
        x = p.reg / p.r0;
    } else {
        if (r > p.cutoff) return 0.0;
        x = fabs(r) / p.r0;
    }
    return -keldysh_pref(p) / p.r0 * keldysh_dxf(x, gl);
}

}  // namespace xatu_dev
