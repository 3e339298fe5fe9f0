// CPU FP64 oracle for the Xatu BSE hot path -- TEST INFRASTRUCTURE ONLY.
//
// A plain restatement of the reference algorithm, element by element, in the
// reference's own loop order, used by tests/ (and bench.py's CPU-baseline leg) as
// the checker of the HIP path. Nothing in the product links or calls it.
//
// Restated functions (reference file:line):
//   oracle_stvh0           src/ExcitonTB.cpp:306-337   (Zhang & Jin STVH0)
//   oracle_potential       src/ExcitonTB.cpp:347-381   (keldysh / coulomb)
//   oracle_truncate_supercell src/Lattice.cpp:293-341
//   oracle_motif_ft        src/ExcitonTB.cpp:438-476   (motifFourierTransform / motifFTMatrix)
//   oracle_ft_stack        src/ExcitonTB.cpp:761-781   (initializeMotifFT loop, OpenMP over mesh)
//   oracle_find_equivalent src/Lattice.cpp:228-279     (findEquivalentPointBZ)
//   oracle_bse             src/ExcitonTB.cpp:520-548, 840-921 (realSpaceInteractionTerm, BShamiltonian)
//   oracle_bse_columns, oracle_bse_apply: the same elements, selected columns and H X
// Extra potential kinds (no reference counterpart) give the explicit derivatives
// dV/dr0 and dV/deps_m used to build dH for the gradient checks.
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

namespace {

constexpr double PI_ = 3.141592653589793;
constexpr double EC_ = 1.6021766E-19;
constexpr double EPS0_ = 8.8541878E-12;

using cd = std::complex<double>;

struct Pot {
    int kind;  // 0 keldysh, 1 coulomb, 2 d(keldysh)/d r0, 3 d(keldysh)/d eps_m (= d/d eps_s)
    double eps_m, eps_s, r0, reg, cutoff;  // cutoff = |a1| * cutoff_ + 1e-5
};

double stvh0(double X) {
    double A0, BY0, P0, Q0, S, T, T2, TA0, R;
    S = 1.0;
    R = 1.0;
    if (X <= 20.0) {
        A0 = 2.0 * X / PI_;
        for (int K = 1; K < 61; K++) {
            R = -R * X / (2.0 * K + 1.0) * X / (2.0 * K + 1.0);
            S = S + R;
            if (std::fabs(R) < std::fabs(S) * 1.0e-12) break;
        }
        return A0 * S;
    }
    int KM = int(0.5 * (X + 1.0));
    if (X >= 50.0) KM = 25;
    for (int K = 1; K <= KM; K++) {
        R = -R * std::pow((2.0 * K - 1.0) / X, 2);
        S = S + R;
        if (std::fabs(R) < std::fabs(S) * 1.0e-12) break;
    }
    T = 4.0 / X;
    T2 = T * T;
    P0 = ((((-.37043e-5 * T2 + .173565e-4) * T2 - .487613e-4) * T2 + .17343e-3) * T2 - 0.1753062e-2) * T2 +
         .3989422793;
    Q0 = T * (((((.32312e-5 * T2 - 0.142078e-4) * T2 + 0.342468e-4) * T2 - 0.869791e-4) * T2 + 0.4564324e-3) * T2 -
              0.0124669441);
    TA0 = X - 0.25 * PI_;
    BY0 = 2.0 / std::sqrt(X) * (P0 * std::sin(TA0) + Q0 * std::cos(TA0));
    return 2.0 / (PI_ * X) * S + BY0;
}

// Struve H1 by its power series (x <= 20).
double h1_series(double x) {
    double s = 1.0, u = 1.0;
    for (int k = 1; k < 80; ++k) {
        u = -u * x * x / ((2.0 * k + 1.0) * (2.0 * k + 3.0));
        s += u;
        if (std::fabs(u) < std::fabs(s) * 1e-16) break;
    }
    return 2.0 * x * x / (3.0 * PI_) * s;
}

double keldysh(double r, const Pot& p) {
    const double eps_bar = (p.eps_m + p.eps_s) / 2;
    const double R = std::fabs(r) / p.r0;
    if (r == 0) return EC_ / (8E-10 * EPS0_ * eps_bar * p.r0) * (stvh0(p.reg / p.r0) - y0(p.reg / p.r0));
    if (r > p.cutoff) return 0.0;
    return EC_ / (8E-10 * EPS0_ * eps_bar * p.r0) * (stvh0(R) - y0(R));
}

double coulomb(double r, const Pot& p) {
    r = std::fabs(r);
    if (r > p.cutoff) return 0.0;
    return (r != 0) ? EC_ / (4E-10 * PI_ * EPS0_ * r) : EC_ * 1E10 / (4 * PI_ * EPS0_ * p.reg);
}

// g(x) = d(x f)/dx = f + x f' for f = STVH0(x) - y0(x), the function the reference's
// Keldysh potential evaluates (no reference counterpart: the r0 derivative of that
// potential, for the gradient checks).
//  x <= 5:      f + x (2/pi - H1 + Y1), H1 by its power series, libm y1;
//  5 < x <= 20: f and x f' cancel to O(x^-3), so g comes from the integral
//               f(x) = (2/pi) int_0^inf e^{-xt} (1 + t^2)^(-1/2) dt, which gives
//               g(x) = (2/pi) int_0^inf e^{-u} (u^2/x^3) (1 + u^2/x^2)^(-3/2) du,
//               summed with the exp-sinh rule u = exp(pi/2 sinh t), step 1/32, t in [-4, 4];
//  x > 20:      the exact derivative of STVH0's asymptotic branch (src/ExcitonTB.cpp:321-336)
//               minus y0: with H0 = 2/(pi x) S + BY0, the S/x terms cancel analytically and
//               g = (2/pi) S'(x) + (BY0 - y0) + x (BY0' + y1), over the same series terms.
double keldysh_dxf(double x) {
    if (x <= 5.0) return stvh0(x) - y0(x) + x * (2.0 / PI_ - h1_series(x) + y1(x));
    if (x <= 20.0) {
        const double h = 1.0 / 32;
        double s = 0.0;
        for (int i = -128; i <= 128; ++i) {
            const double t = i * h;
            const double u = std::exp(0.5 * PI_ * std::sinh(t));
            const double q = 1.0 + (u / x) * (u / x);
            s += std::exp(-u) * u * u / (q * std::sqrt(q)) * u * 0.5 * PI_ * std::cosh(t);
        }
        return 2.0 / PI_ * h * s / (x * x * x);
    }
    int km = int(0.5 * (x + 1.0));
    if (x >= 50.0) km = 25;
    double S = 1.0, R = 1.0, dS = 0.0;
    for (int k = 1; k <= km; ++k) {
        R = -R * std::pow((2.0 * k - 1.0) / x, 2);  // R_k = (-1)^k ((2k-1)!!)^2 x^-2k
        S += R;
        dS += -2.0 * k * R / x;
        if (std::fabs(R) < std::fabs(S) * 1.0e-12) break;
    }
    // BY0 = 2 x^-1/2 (P0(t^2) sin th + t Qp(t^2) cos th), t = 4/x, th = x - pi/4
    const double t = 4.0 / x, t2 = t * t, dt = -t * t / 4.0;
    const double P = ((((-.37043e-5 * t2 + .173565e-4) * t2 - .487613e-4) * t2 + .17343e-3) * t2 - 0.1753062e-2) * t2 +
                     .3989422793;
    const double dP = (((5 * -.37043e-5 * t2 + 4 * .173565e-4) * t2 - 3 * .487613e-4) * t2 + 2 * .17343e-3) * t2 -
                      0.1753062e-2;  // dP/d(t^2)
    const double Qp = ((((.32312e-5 * t2 - 0.142078e-4) * t2 + 0.342468e-4) * t2 - 0.869791e-4) * t2 + 0.4564324e-3) * t2 -
                      0.0124669441;
    const double dQp = (((5 * .32312e-5 * t2 - 4 * 0.142078e-4) * t2 + 3 * 0.342468e-4) * t2 - 2 * 0.869791e-4) * t2 +
                       0.4564324e-3;  // dQp/d(t^2)
    const double Q = t * Qp;
    const double dPx = dP * 2.0 * t * dt;                  // dP0/dx
    const double dQx = (Qp + t * dQp * 2.0 * t) * dt;      // dQ0/dx
    const double th = x - 0.25 * PI_, sn = std::sin(th), cs = std::cos(th);
    const double rx = 1.0 / std::sqrt(x);
    const double by0 = 2.0 * rx * (P * sn + Q * cs);
    const double dby0 = -rx / x * (P * sn + Q * cs) + 2.0 * rx * (dPx * sn + P * cs + dQx * cs - Q * sn);
    return 2.0 / PI_ * dS + (by0 - y0(x)) + x * (dby0 + y1(x));
}

double keldysh_dr0(double r, const Pot& p) {
    const double eps_bar = (p.eps_m + p.eps_s) / 2;
    double x;
    if (r == 0) {
        x = p.reg / p.r0;
    } else {
        if (r > p.cutoff) return 0.0;
        x = std::fabs(r) / p.r0;
    }
    const double C = EC_ / (8E-10 * EPS0_ * eps_bar);
    return -C / (p.r0 * p.r0) * keldysh_dxf(x);
}

double potential(double r, const Pot& p) {
    switch (p.kind) {
        case 0:
            return keldysh(r, p);
        case 1:
            return coulomb(r, p);
        case 2:
            return keldysh_dr0(r, p);
        case 3:
            return -keldysh(r, p) / (2.0 * ((p.eps_m + p.eps_s) / 2));
        default:
            return 0.0;
    }
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// V_ab(k) for all a <= b, lower triangle by conjugation (motifFTMatrix).
void motif_ft_matrix(int natoms, const double* motif, int nR, const double* cells, const double* k, const Pot& p,
                     double total_cells, cd* out) {
    const cd imag(0, 1);
    for (int a = 0; a < natoms; ++a) {
        for (int b = a; b < natoms; ++b) {
            cd Vk = 0.0;
            for (int n = 0; n < nR; ++n) {
                const double* cell = cells + 3 * n;
                double v[3];
                for (int x = 0; x < 3; ++x) v[x] = cell[x] + motif[3 * a + x] - motif[3 * b + x];
                const double module = norm3(v);
                const double dot = k[0] * cell[0] + k[1] * cell[1] + k[2] * cell[2];
                Vk += potential(module, p) * std::exp(imag * dot);
            }
            Vk /= total_cells;
            out[a + natoms * b] = Vk;
            out[b + natoms * a] = std::conj(Vk);
        }
    }
}

}  // namespace

extern "C" {

double oracle_stvh0(double x) { return stvh0(x); }
double oracle_y0(double x) { return y0(x); }
double oracle_h1(double x) { return h1_series(x); }
double oracle_keldysh_dxf(double x) { return keldysh_dxf(x); }

double oracle_potential(int kind, double r, double eps_m, double eps_s, double r0, double reg, double cutoff) {
    Pot p{kind, eps_m, eps_s, r0, reg, cutoff};
    return potential(r, p);
}

// Cells (Cartesian, nR x 3) of the centred ncell^ndim supercell with |R| < radius + 1e-5.
// Returns nR; writes at most max_cells rows.
int oracle_truncate_supercell(int ndim, const double* bravais, int ncell, double radius, double* out,
                              int max_cells) {
    long long total = 1;
    for (int j = 0; j < ndim; ++j) total *= ncell;
    const int shift = ncell / 2;
    int count = 0;
    for (long long i = 0; i < total; ++i) {
        long long rem = i;
        double v[3] = {0, 0, 0};
        for (int j = 0; j < ndim; ++j) {
            const double m = static_cast<double>(rem % ncell) - shift;
            rem /= ncell;
            for (int x = 0; x < 3; ++x) v[x] += m * bravais[3 * j + x];
        }
        if (norm3(v) < radius + 1E-5) {
            if (count < max_cells && out) {
                for (int x = 0; x < 3; ++x) out[3 * count + x] = v[x];
            }
            ++count;
        }
    }
    return count;
}

void oracle_motif_ft(int natoms, const double* motif, int nR, const double* cells, const double* k, int kind,
                     double eps_m, double eps_s, double r0, double reg, double cutoff, double total_cells,
                     double* out) {
    Pot p{kind, eps_m, eps_s, r0, reg, cutoff};
    motif_ft_matrix(natoms, motif, nR, cells, k, p, total_cells, reinterpret_cast<cd*>(out));
}

// ftMotifStack for every mesh point (natoms x natoms x nmesh, column-major complex).
void oracle_ft_stack(int natoms, const double* motif, int nR, const double* cells, int nmesh, const double* mesh,
                     int kind, double eps_m, double eps_s, double r0, double reg, double cutoff, double total_cells,
                     double* out) {
    Pot p{kind, eps_m, eps_s, r0, reg, cutoff};
    cd* o = reinterpret_cast<cd*>(out);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nmesh; ++i) {
        motif_ft_matrix(natoms, motif, nR, cells, mesh + 3 * static_cast<size_t>(i), p, total_cells,
                        o + static_cast<size_t>(i) * natoms * natoms);
    }
}

// findEquivalentPointBZ with the inverse of (B B^T) given (ndim x ndim row-major).
int oracle_find_equivalent(int ndim, const double* recip, const double* inv, const double* kpoint, int ncell,
                           int factor) {
    ncell = ncell * factor;
    double indep[3] = {0, 0, 0}, coefs[3] = {0, 0, 0};
    for (int i = 0; i < ndim; ++i)
        for (int x = 0; x < 3; ++x) indep[i] += recip[3 * i + x] * kpoint[x];
    for (int i = 0; i < ndim; ++i) {
        double s = 0;
        for (int j = 0; j < ndim; ++j) s += inv[i * ndim + j] * indep[j];
        coefs[i] = s * 2 * ncell;
    }
    int index = 0;
    const int cells_array[3] = {1, ncell, ncell * ncell};
    for (int i = 0; i < ndim; ++i) {
        double c = (ncell % 2 == 1) ? coefs[i] - 1 : coefs[i];
        c = std::round(c);
        if (c >= ncell) {
            c -= 2 * ncell;
        } else if (c < -ncell) {
            c += 2 * ncell;
        }
        c += ncell;
        c /= 2;
        index += static_cast<int>(c * cells_array[i]);
    }
    return index;
}

// Full Bethe-Salpeter matrix (HBS + HBS^H) for an explicit basis.
//  basis: N x 3 {vcol, ccol, k} (stack columns, k index)
//  kpoints: nk x 3; stacks: eigvec norb x nb x nk complex, eigval nb x nk
//  orb_atom: atom of every orbital (norb)
//  ft: natoms x natoms x nmesh complex, ftq: natoms x natoms
//  out: N x N complex, column-major
// The reference's BSE element rule (BShamiltonian, src/ExcitonTB.cpp:857-919) for the
// upper triangle i <= j of the basis rows {v, c, k}: -(D - X) off the diagonal and
// h + conj(h) with h = (scissor + Ec - Ev)/2 - (D - X)/2 on it, D and X from
// realSpaceInteractionTerm (src/ExcitonTB.cpp:520-548: atom-resolved densities contracted
// with the lattice-sum matrix by arma::dot, no conjugation).
struct BseElements {
    const int* basis;
    const double* kpoints;
    int norb, nb, natoms;
    const cd *EK, *EKQ, *FT, *FTQ;
    const double *ek, *ekq;
    const int* orb_atom;
    int exchange;
    double scissor;
    int ndim;
    const double *recip, *inv;
    int ncell, factor;

    cd term(const cd* c1, const cd* c2, const cd* c3, const cd* c4, const cd* M) const {
        cd r1[16], r2[16];
        for (int a = 0; a < natoms; ++a) r1[a] = r2[a] = 0.0;
        for (int al = 0; al < norb; ++al) {
            r1[orb_atom[al]] += std::conj(c1[al]) * c3[al];
            r2[orb_atom[al]] += std::conj(c2[al]) * c4[al];
        }
        cd t = 0.0;
        for (int a = 0; a < natoms; ++a) {
            cd y = 0.0;
            for (int b = 0; b < natoms; ++b) y += M[a + natoms * b] * r2[b];
            t += r1[a] * y;  // arma::dot: no conjugation
        }
        return t;
    }

    cd upper(long long i, long long j) const {
        const int v = basis[3 * i], c = basis[3 * i + 1], k = basis[3 * i + 2];
        const int v2 = basis[3 * j], c2 = basis[3 * j + 1], k2 = basis[3 * j + 2];
        const cd* coefsK = EK + (static_cast<size_t>(k) * nb + v) * norb;
        const cd* coefsKQ = EKQ + (static_cast<size_t>(k) * nb + c) * norb;
        const cd* coefsK2 = EK + (static_cast<size_t>(k2) * nb + v2) * norb;
        const cd* coefsK2Q = EKQ + (static_cast<size_t>(k2) * nb + c2) * norb;
        double dk[3];
        for (int x = 0; x < 3; ++x) dk[x] = kpoints[3 * k2 + x] - kpoints[3 * k + x];
        const int q = oracle_find_equivalent(ndim, recip, inv, dk, ncell, factor);
        const cd D = term(coefsKQ, coefsK2, coefsK2Q, coefsK, FT + static_cast<size_t>(q) * natoms * natoms);
        cd X = 0.0;
        if (exchange) X = term(coefsKQ, coefsK2, coefsK, coefsK2Q, FTQ);
        if (i == j) {
            const cd h = (scissor + ekq[static_cast<size_t>(k) * nb + c] - ek[static_cast<size_t>(k) * nb + v]) / 2. -
                         (D - X) / 2.;
            return h + std::conj(h);
        }
        return -(D - X);
    }

    // H(i, j) of the Hermitian matrix HBS + HBS^H
    cd at(long long i, long long j) const { return i <= j ? upper(i, j) : std::conj(upper(j, i)); }
};

#define XATU_ORACLE_BSE_ARGS                                                                                 \
    int N, const int *basis, int nk, const double *kpoints, int norb, int nb, const double *evk,             \
        const double *evkq, const double *ek, const double *ekq, int natoms, const int *orb_atom,            \
        const double *ft, const double *ftq, int exchange, double scissor, int ndim, const double *recip,    \
        const double *inv, int ncell, int factor
#define XATU_ORACLE_BSE_INIT                                                                                  \
    BseElements E{basis, kpoints, norb, nb, natoms, reinterpret_cast<const cd*>(evk),                        \
                  reinterpret_cast<const cd*>(evkq), reinterpret_cast<const cd*>(ft),                        \
                  reinterpret_cast<const cd*>(ftq), ek, ekq, orb_atom, exchange, scissor, ndim, recip, inv,  \
                  ncell, factor};                                                                            \
    (void)nk;

// The whole N x N matrix, column-major.
void oracle_bse(XATU_ORACLE_BSE_ARGS, double* out) {
    XATU_ORACLE_BSE_INIT
    cd* H = reinterpret_cast<cd*>(out);
    const uint64_t n = static_cast<uint64_t>(N);
#pragma omp parallel for schedule(dynamic, 64)
    for (long long i = 0; i < N; ++i) {
        for (long long j = i; j < N; ++j) {
            const cd h = E.upper(i, j);
            H[i + j * n] = h;
            if (i != j) H[j + i * n] = std::conj(h);
        }
    }
}

// Columns cols[0..ncols) of the matrix (N x ncols, column-major).
void oracle_bse_columns(XATU_ORACLE_BSE_ARGS, int ncols, const int* cols, double* out) {
    XATU_ORACLE_BSE_INIT
    cd* O = reinterpret_cast<cd*>(out);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < N; ++i)
        for (int c = 0; c < ncols; ++c) O[i + static_cast<size_t>(c) * N] = E.at(i, cols[c]);
}

// Y = H X for nvec vectors (N x nvec, column-major) without storing H: every element is
// evaluated by the element rule for its row.
void oracle_bse_apply(XATU_ORACLE_BSE_ARGS, int nvec, const double* x, double* y) {
    XATU_ORACLE_BSE_INIT
    const cd* X = reinterpret_cast<const cd*>(x);
    cd* Y = reinterpret_cast<cd*>(y);
#pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < N; ++i) {
        std::vector<cd> acc(nvec, 0.0);
        for (long long j = 0; j < N; ++j) {
            const cd h = E.at(i, j);
            for (int n = 0; n < nvec; ++n) acc[n] += h * X[j + static_cast<size_t>(n) * N];
        }
        for (int n = 0; n < nvec; ++n) Y[i + static_cast<size_t>(n) * N] = acc[n];
    }
}

}  // extern "C"
