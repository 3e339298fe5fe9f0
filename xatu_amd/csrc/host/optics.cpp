// Optical response of the exciton states: single-particle velocity matrix elements,
// exciton oscillator strengths and the Lorentzian-broadened exciton conductivity
// (reference: ResultTB::excitonOscillatorStrength and writeAbsorptionSpectrum,
// src/ResultTB.cpp:242-298, 412-460, which call the Fortran routines of
// src/skubo_w.f90). The velocity matrix elements of the band stacks are computed here
// on the host, like the band problem; the contraction with the exciton eigenvectors
// runs on the GPU (xatu_oscillator_strengths).
#include <cmath>
#include <stdexcept>

#include "exciton.hpp"
#include "xatu_bse.h"

namespace xatu {

namespace {
constexpr double kBohr = 0.52917721067121;  // Angstrom (skubo_w.f90:229-237)
constexpr double kHartree = 27.211385;      // eV
}  // namespace

// V[((k*3 + j)*nb + n)*nb + m], atomic units (src/skubo_w.f90:175-330 with
// hoppings_observables_TB :407-441, get_vme_kernels :445-527, get_eigen_vme :531-627).
// The reference's construction is kept: the kernels are summed over the lower orbital
// triangle and mirrored (the orbital diagonal of the derivative kernel ends up
// conjugated), the position operator uses the motif at the first listed cell only,
// the overlap is the identity at every cell, the z components of R are dropped, and
// the band diagonal of V is conjugated by the same mirroring.
std::vector<cdouble> velocityMatrixElements(const ExcitonTB& ex) {
    const SystemTB& s = *ex.system;
    if (!s.overlapMatrices.empty()) {
        throw std::invalid_argument("velocity matrix elements need an orthonormal basis (no overlap matrices)");
    }
    const int norb = s.basisdim, nR = s.ncells, nk = s.nk;
    const int nb = static_cast<int>(ex.bandList.size());
    const cdouble I(0.0, 1.0);
    std::vector<double> B(static_cast<size_t>(norb) * 3);  // orbital positions (bohr)
    {
        const std::vector<int> opa = s.orbitalsPerAtom();
        int it = 0;
        for (int a = 0; a < s.natoms; ++a)
            for (int o = 0; o < opa[a]; ++o, ++it)
                for (int x = 0; x < 3; ++x) B[3 * it + x] = s.motif[4 * a + x] / kBohr;
    }
    std::vector<cdouble> V(static_cast<size_t>(nk) * 3 * nb * nb);
#pragma omp parallel for schedule(dynamic, 8)
    for (int k = 0; k < nk; ++k) {
        const double kx = s.kpoints[3 * static_cast<size_t>(k)] * kBohr;
        const double ky = s.kpoints[3 * static_cast<size_t>(k) + 1] * kBohr;
        std::vector<cdouble> f(nR), sderR(static_cast<size_t>(nR) * 2);
        for (int r = 0; r < nR; ++r) {
            const double Rx = s.unitCellList[3 * static_cast<size_t>(r)] / kBohr;
            const double Ry = s.unitCellList[3 * static_cast<size_t>(r) + 1] / kBohr;
            f[r] = std::exp(I * (kx * Rx + ky * Ry));
            sderR[2 * r] = I * Rx;
            sderR[2 * r + 1] = I * Ry;
        }
        // lower-triangle kernels, then the mirrored full matrices
        std::vector<cdouble> hder(3 * static_cast<size_t>(norb) * norb, 0.0), aker(hder.size(), 0.0);
        auto at = [norb](int j, int a, int b) { return (static_cast<size_t>(j) * norb + b) * norb + a; };
        cdouble sdiag[3] = {0.0, 0.0, 0.0};  // sum_R f(R) i R_j (the identity overlap)
        for (int r = 0; r < nR; ++r)
            for (int j = 0; j < 2; ++j) sdiag[j] += f[r] * sderR[2 * r + j];
        for (int a = 0; a < norb; ++a) {
            for (int b = 0; b <= a; ++b) {
                for (int j = 0; j < 3; ++j) {
                    cdouble h = 0.0;
                    if (j < 2) {
                        for (int r = 0; r < nR; ++r) {
                            const cdouble hr = s.hamiltonianMatrices[(static_cast<size_t>(r) * norb + b) * norb + a] /
                                               kHartree;
                            h += f[r] * sderR[2 * r + j] * hr;
                        }
                    }
                    const cdouble sd = (a == b) ? sdiag[j] : cdouble(0.0);
                    const cdouble ak = (a == b ? f[0] * B[3 * a + j] : cdouble(0.0)) + I * sd;
                    // lower entry, then its mirror (for a == b the mirror overwrites it)
                    hder[at(j, a, b)] = h;
                    aker[at(j, a, b)] = ak;
                    hder[at(j, b, a)] = std::conj(h);
                    aker[at(j, b, a)] = std::conj(ak) + I * std::conj(sd);
                }
            }
        }
        const cdouble* C = ex.eigvecKStack.data() + static_cast<size_t>(k) * nb * norb;
        const double* E = ex.eigvalKStack.data() + static_cast<size_t>(k) * nb;
        for (int j = 0; j < 3; ++j) {
            for (int n = 0; n < nb; ++n) {
                for (int m = 0; m <= n; ++m) {
                    cdouble va = 0.0, t1 = 0.0, t2 = 0.0;
                    for (int a = 0; a < norb; ++a) {
                        const cdouble cn = std::conj(C[static_cast<size_t>(n) * norb + a]);
                        for (int b = 0; b < norb; ++b) {
                            const cdouble w = cn * C[static_cast<size_t>(m) * norb + b];
                            va += w * hder[at(j, a, b)];
                            t1 += w * aker[at(j, a, b)];
                            t2 += w * std::conj(aker[at(j, b, a)]);
                        }
                    }
                    const cdouble v = va + I * (E[n] / kHartree * t1 - E[m] / kHartree * t2);
                    V[((static_cast<size_t>(k) * 3 + j) * nb + n) * nb + m] = v;
                    V[((static_cast<size_t>(k) * 3 + j) * nb + m) * nb + n] = std::conj(v);
                }
            }
        }
    }
    return V;
}

// vme_ex(j, n) = sum_{k,c,v} X_n(kcv) V[k, j, v, nv + c] (skubo_w.f90:300-323): 3 x ncols.
std::vector<cdouble> ResultTB::excitonOscillatorStrength() const {
    const ExcitonTB& ex = *exciton;
    const int nb = static_cast<int>(ex.bandList.size());
    const int nv = static_cast<int>(ex.valenceBands.size());
    const int nc = static_cast<int>(ex.conductionBands.size());
    const long long N = ex.excitonbasisdim;
    const std::vector<cdouble> V = velocityMatrixElements(ex);
    // pair velocities vel[i + N j] for basis row i = k nc nv + c nv + v
    std::vector<cdouble> vel(static_cast<size_t>(N) * 3);
    for (int k = 0; k < ex.system->nk; ++k)
        for (int c = 0; c < nc; ++c)
            for (int v = 0; v < nv; ++v)
                for (int j = 0; j < 3; ++j)
                    vel[static_cast<size_t>(k) * nc * nv + c * nv + v + static_cast<size_t>(N) * j] =
                        V[((static_cast<size_t>(k) * 3 + j) * nb + v) * nb + nv + c];
    std::vector<cdouble> out(static_cast<size_t>(3) * ncols);
    ex.check(xatu_oscillator_strengths(ex.context(), ncols, reinterpret_cast<const double*>(eigvec.data()),
                                       reinterpret_cast<const double*>(vel.data()),
                                       reinterpret_cast<double*>(out.data())));
    return out;
}

// Exciton conductivity sigma_jj'(w), j, j' in {x, y}, broadened with Lorentzians of
// width eta (eV) on nw frequencies w0 + wrange/nw * i (eV), in the reference's
// units (skubo_w.f90:95-130 and broad_vector :684-720): returns nw x 5 values
// {w, s_xx, s_xy, s_yx, s_yy}.
std::vector<double> ResultTB::absorptionSpectrum(double w0, double wrange, int nw, double eta) const {
    if (nw < 1 || eta <= 0.0) throw std::invalid_argument("absorptionSpectrum: nw >= 1 and eta > 0 required");
    const SystemTB& s = *exciton->system;
    const std::vector<cdouble> vme = excitonOscillatorStrength();
    double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < s.ndim; ++i)
        for (int x = 0; x < 3; ++x) R[i][x] = s.bravaisLattice[3 * i + x] / kBohr;
    const double cx = R[0][1] * R[1][2] - R[0][2] * R[1][1];
    const double cy = R[0][2] * R[1][0] - R[0][0] * R[1][2];
    const double cz = R[0][0] * R[1][1] - R[0][1] * R[1][0];
    const double vcell = std::sqrt(cx * cx + cy * cy + cz * cz);
    const double pre = M_PI / (static_cast<double>(s.nk) * vcell);
    const double et = eta / kHartree;
    std::vector<double> out(static_cast<size_t>(nw) * 5, 0.0);
    for (int i = 0; i < nw; ++i) {
        const double w = (w0 + wrange / nw * i) / kHartree;
        out[5 * static_cast<size_t>(i)] = w * kHartree;
        for (int n = 0; n < ncols; ++n) {
            const double e = eigval[n] / kHartree;
            const double L = (1.0 / M_PI) * (et / ((w - e) * (w - e) + et * et));
            for (int j = 0; j < 2; ++j)
                for (int jp = 0; jp < 2; ++jp) {
                    const cdouble sk = pre * std::conj(vme[3 * static_cast<size_t>(n) + j]) *
                                       vme[3 * static_cast<size_t>(n) + jp] / e;
                    out[5 * static_cast<size_t>(i) + 1 + 2 * j + jp] += sk.real() * L;
                }
        }
    }
    return out;
}

}  // namespace xatu
