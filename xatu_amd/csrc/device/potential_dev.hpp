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
// u_k = -u_{k-1} x^2 / ((2k+1)(2k+3)); valid (to ~1e-13) for x <= 20.
// For x > 20 the asymptotic expansion of H1 - Y1 is used by the caller.
__device__ inline double struve_h1_series(double x) {
    double s = 1.0, u = 1.0;
    for (int k = 1; k < 80; ++k) {
        u = -u * x * x / ((2.0 * k + 1.0) * (2.0 * k + 3.0));
        s += u;
        if (fabs(u) < fabs(s) * 1.0e-16) break;
    }
    return 2.0 * x * x / (3.0 * kPi) * s;
}

// dV/dr0 for the Keldysh potential at distance r (same special cases as V):
// with f = H0 - Y0 and x = r/r0, V = C f(x) / (eps_bar r0), so
// dV/dr0 = -C / (eps_bar r0^2) * (f(x) + x f'(x)),  f'(x) = 2/pi - H1(x) + Y1(x).
__device__ inline double keldysh_dr0(double r, const PotDesc& p) {
    double x;
    if (r == 0.0) {
        x = p.reg / p.r0;
    } else {
        if (r > p.cutoff) return 0.0;
        x = fabs(r) / p.r0;
    }
    const double f = struve_h0(x) - y0(x);
    double fp;
    if (x <= 20.0) {
        fp = 2.0 / kPi - struve_h1_series(x) + y1(x);
    } else {
        // H1 - Y1 ~ (2/pi) (1 + 1/x^2 - 3/x^4 + 45/x^6)
        const double ix2 = 1.0 / (x * x);
        fp = 2.0 / kPi - (2.0 / kPi) * (1.0 + ix2 * (1.0 + ix2 * (-3.0 + ix2 * 45.0)));
    }
    return -keldysh_pref(p) / p.r0 * (f + x * fp);
}

}  // namespace xatu_dev
