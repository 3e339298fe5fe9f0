// ResultTB observables and writers beyond the eigen-data (reference:
// include/xatu/Result.tpp:66-160, 196-340, 343-402, 456-500; src/ResultTB.cpp:21-99,
// 310-411, 470-482). They post-process the converged states on the host; the state
// vectors use the exciton basis order row = k (nc nv) + c nv + v.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>

#include "exciton.hpp"

namespace xatu {

namespace {

std::vector<cdouble> state_column(const ResultTB& r, int stateindex, const char* who) {
    if (stateindex < 0 || stateindex >= r.ncols) {
        throw std::invalid_argument(std::string(who) + ": state index out of range");
    }
    return std::vector<cdouble>(r.eigvec.begin() + static_cast<size_t>(stateindex) * r.nrows,
                                r.eigvec.begin() + static_cast<size_t>(stateindex + 1) * r.nrows);
}

void check_state(const ResultTB& r, const std::vector<cdouble>& coefs, const char* who) {
    if (static_cast<long long>(coefs.size()) != r.nrows) {
        throw std::invalid_argument(std::string(who) + ": the state must have one coefficient per basis row");
    }
}

}  // namespace

// <S_z> of the hole, the electron and their sum (src/ResultTB.cpp:21-99). The spin of
// orbital o of an atom is +1/2 for even o and -1/2 for odd o (spin-up and spin-down
// orbitals alternate within each atom). As the reference builds them, the hole matrix
// at k is S^h(i, j) = sum_a conj(V_j(a)) s_a V_i(a) over the valence bands of the k
// stack and the electron matrix S^e(i, j) = sum_a conj(C_i(a)) s_a C_j(a) over the
// conduction bands of the k+Q stack; hole = -psi^H (1 (x) S^h) psi, electron =
// psi^H (S^e (x) 1) psi, total = Re(hole + electron).
std::vector<cdouble> ResultTB::spinX(const std::vector<cdouble>& coefs) const {
    check_state(*this, coefs, "spinX");
    const SystemTB& s = *exciton->system;
    const std::vector<int> opa = s.orbitalsPerAtom();
    std::vector<double> sz;
    for (int n : opa) {
        if (n % 2 != 0) throw std::logic_error("spinX: every atom needs pairs of spin orbitals");
        for (int o = 0; o < n; ++o) sz.push_back((o % 2 == 0) ? 0.5 : -0.5);
    }
    const int bd = s.basisdim;
    const int nb = static_cast<int>(exciton->bandList.size());
    const int nv = static_cast<int>(exciton->valenceBands.size());
    const int nc = static_cast<int>(exciton->conductionBands.size());
    std::vector<int> vidx(nv), cidx(nc);
    for (int i = 0; i < nv; ++i) vidx[i] = exciton->bandToIndex.at(static_cast<int>(exciton->valenceBands[i]));
    for (int i = 0; i < nc; ++i) cidx[i] = exciton->bandToIndex.at(static_cast<int>(exciton->conductionBands[i]));
    cdouble hole(0.0, 0.0), electron(0.0, 0.0);
    std::vector<cdouble> Sh(static_cast<size_t>(nv) * nv), Se(static_cast<size_t>(nc) * nc);
    for (int k = 0; k < s.nk; ++k) {
        auto vk = [&](int col) { return exciton->eigvecKStack.data() + (static_cast<size_t>(k) * nb + col) * bd; };
        auto ck = [&](int col) { return exciton->eigvecKQStack.data() + (static_cast<size_t>(k) * nb + col) * bd; };
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nv; ++j) {
                const cdouble* a = vk(vidx[j]);
                const cdouble* b = vk(vidx[i]);
                cdouble t(0.0, 0.0);
                for (int o = 0; o < bd; ++o) t += std::conj(a[o]) * (b[o] * sz[o]);
                Sh[i + static_cast<size_t>(j) * nv] = t;
            }
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j) {
                const cdouble* a = ck(cidx[i]);
                const cdouble* b = ck(cidx[j]);
                cdouble t(0.0, 0.0);
                for (int o = 0; o < bd; ++o) t += std::conj(a[o]) * (b[o] * sz[o]);
                Se[i + static_cast<size_t>(j) * nc] = t;
            }
        const cdouble* x = coefs.data() + static_cast<size_t>(k) * nc * nv;
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < nv; ++i)
                for (int j = 0; j < nv; ++j) hole -= std::conj(x[c * nv + i]) * Sh[i + static_cast<size_t>(j) * nv] * x[c * nv + j];
        for (int v = 0; v < nv; ++v)
            for (int i = 0; i < nc; ++i)
                for (int j = 0; j < nc; ++j) electron += std::conj(x[i * nv + v]) * Se[i + static_cast<size_t>(j) * nc] * x[j * nv + v];
    }
    return {cdouble((hole + electron).real(), 0.0), hole, electron};
}

std::vector<cdouble> ResultTB::spinX(int stateindex) const { return spinX(state_column(*this, stateindex, "spinX")); }

// Result.tpp:456-470
void ResultTB::writeSpin(int n, FILE* f) const {
    if (n > nrows || n < 0) {
        throw std::invalid_argument("Optional argument n must be a positive integer equal or below basisdim");
    }
    const int m = (n == 0) ? ncols : std::min(n, ncols);
    fprintf(f, "n\tSt\tSe\tSh\n");
    for (int i = 0; i < m; ++i) {
        const std::vector<cdouble> sp = spinX(i);
        fprintf(f, "%d\t%11.7lf\t%11.7lf\t%11.7lf\n", i, sp[0].real(), sp[1].real(), sp[2].real());
    }
}

// Probability of the electron on atom e of cell eCell with the hole on atom h of cell
// hCell (src/ResultTB.cpp:310-368, 470-482): with R = eCell - hCell,
//   sum_{alpha in e, beta in h} | sum_{kcv} C_kc(alpha) conj(V_kv(beta)) e^{i k.R} psi_kcv |^2
// (C from the k+Q stack, V from the k stack; no conjugation of psi).
double ResultTB::realSpaceWavefunction(const std::vector<cdouble>& coefs, int electronIndex, int holeIndex,
                                       const std::vector<double>& eCell, const std::vector<double>& hCell) const {
    check_state(*this, coefs, "realSpaceWavefunction");
    const SystemTB& s = *exciton->system;
    if (electronIndex < 0 || electronIndex >= s.natoms || holeIndex < 0 || holeIndex >= s.natoms) {
        throw std::invalid_argument("realSpaceWavefunction: atom index out of range");
    }
    if (eCell.size() != 3 || hCell.size() != 3) throw std::invalid_argument("realSpaceWavefunction: cells are 3-vectors");
    const std::vector<int> opa = s.orbitalsPerAtom();
    int eIndex = 0, hIndex = 0;
    for (int a = 0; a < electronIndex; ++a) eIndex += opa[a];
    for (int a = 0; a < holeIndex; ++a) hIndex += opa[a];
    const int bd = s.basisdim;
    const int nb = static_cast<int>(exciton->bandList.size());
    const int nv = static_cast<int>(exciton->valenceBands.size());
    const int nc = static_cast<int>(exciton->conductionBands.size());
    std::vector<int> vidx(nv), cidx(nc);
    for (int i = 0; i < nv; ++i) vidx[i] = exciton->bandToIndex.at(static_cast<int>(exciton->valenceBands[i]));
    for (int i = 0; i < nc; ++i) cidx[i] = exciton->bandToIndex.at(static_cast<int>(exciton->conductionBands[i]));
    const double R[3] = {eCell[0] - hCell[0], eCell[1] - hCell[1], eCell[2] - hCell[2]};
    std::vector<cdouble> phased(coefs.size());
    for (int k = 0; k < s.nk; ++k) {
        const double* kp = s.kpoints.data() + 3 * static_cast<size_t>(k);
        const double ph = kp[0] * R[0] + kp[1] * R[1] + kp[2] * R[2];
        const cdouble e(std::cos(ph), std::sin(ph));
        for (int b = 0; b < nc * nv; ++b) phased[static_cast<size_t>(k) * nc * nv + b] = coefs[static_cast<size_t>(k) * nc * nv + b] * e;
    }
    double total = 0.0;
    for (int al = 0; al < opa[electronIndex]; ++al)
        for (int be = 0; be < opa[holeIndex]; ++be) {
            cdouble acc(0.0, 0.0);
            for (int k = 0; k < s.nk; ++k) {
                const cdouble* ev = exciton->eigvecKStack.data() + static_cast<size_t>(k) * nb * bd;
                const cdouble* ec = exciton->eigvecKQStack.data() + static_cast<size_t>(k) * nb * bd;
                for (int c = 0; c < nc; ++c) {
                    const cdouble cv = ec[static_cast<size_t>(cidx[c]) * bd + eIndex + al];
                    for (int v = 0; v < nv; ++v) {
                        const cdouble vv = ev[static_cast<size_t>(vidx[v]) * bd + hIndex + be];
                        acc += cv * std::conj(vv) * phased[static_cast<size_t>(k) * nc * nv + c * nv + v];
                    }
                }
            }
            total += std::norm(acc);
        }
    return total;
}

// Result.tpp:343-376: the hole position, then one line per (cell, atom) of the cells of
// truncateSupercell(ncell, |a1| ncells), and "#".
void ResultTB::writeRealspaceAmplitude(const std::vector<cdouble>& coefs, int holeIndex,
                                       const std::vector<double>& holeCell, FILE* f, int ncells) const {
    const SystemTB& s = *exciton->system;
    if (holeIndex < 0 || holeIndex >= s.natoms) throw std::invalid_argument("writeRealspaceAmplitude: bad hole index");
    if (holeCell.size() != 3) throw std::invalid_argument("writeRealspaceAmplitude: the hole cell is a 3-vector");
    const double* tau_h = s.motif.data() + 4 * static_cast<size_t>(holeIndex);
    fprintf(f, "%11.8lf\t%11.8lf\t%14.11lf\n", tau_h[0] + holeCell[0], tau_h[1] + holeCell[1], 0.0);
    const double* a1 = s.bravaisLattice.data();
    const double radius = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]) * ncells;
    const std::vector<double> cells = s.truncateSupercell(exciton->ncell, radius);
    const size_t ncell_rows = cells.size() / 3;
    std::vector<double> amp(ncell_rows * s.natoms);
#pragma omp parallel for schedule(dynamic)
    for (long long ci = 0; ci < static_cast<long long>(ncell_rows); ++ci) {
        const std::vector<double> cell(cells.begin() + 3 * ci, cells.begin() + 3 * ci + 3);
        for (int a = 0; a < s.natoms; ++a) amp[ci * s.natoms + a] = realSpaceWavefunction(coefs, a, holeIndex, cell, holeCell);
    }
    for (size_t ci = 0; ci < ncell_rows; ++ci)
        for (int a = 0; a < s.natoms; ++a) {
            const double* tau = s.motif.data() + 4 * static_cast<size_t>(a);
            fprintf(f, "%11.8lf\t%11.8lf\t%14.11lf\n", tau[0] + cells[3 * ci], tau[1] + cells[3 * ci + 1],
                    amp[ci * s.natoms + a]);
        }
    fprintf(f, "#\n");
}

void ResultTB::writeRealspaceAmplitude(int stateindex, int holeIndex, const std::vector<double>& holeCell, FILE* f,
                                       int ncells) const {
    writeRealspaceAmplitude(state_column(*this, stateindex, "writeRealspaceAmplitude"), holeIndex, holeCell, f, ncells);
}

// Result.tpp:479-485: k index of the largest |coefficient| of the state.
int ResultTB::findExcitonPeak(int stateindex) const {
    const std::vector<cdouble> c = state_column(*this, stateindex, "findExcitonPeak");
    size_t best = 0;
    for (size_t i = 1; i < c.size(); ++i)
        if (std::abs(c[i]) > std::abs(c[best])) best = i;
    const long long nbc = static_cast<long long>(exciton->valenceBands.size() * exciton->conductionBands.size());
    return static_cast<int>(static_cast<long long>(best) / nbc);
}

// Result.tpp:86-96: gap between the lowest conduction and highest valence band of the
// exciton, both from the k stack, at the peak of the ground state.
double ResultTB::determineGap() const {
    const int k = findExcitonPeak(0);
    long long vmax = exciton->valenceBands.front(), cmin = exciton->conductionBands.front();
    for (long long b : exciton->valenceBands) vmax = std::max(vmax, b);
    for (long long b : exciton->conductionBands) cmin = std::min(cmin, b);
    const int nb = static_cast<int>(exciton->bandList.size());
    const int vi = exciton->bandToIndex.at(static_cast<int>(vmax));
    const int ci = exciton->bandToIndex.at(static_cast<int>(cmin));
    return exciton->eigvalKStack[static_cast<size_t>(k) * nb + ci] - exciton->eigvalKStack[static_cast<size_t>(k) * nb + vi];
}

// Result.tpp:66-78 (a gap of -1 selects determineGap; another negative gap is reported
// and used as given, as the reference does)
double ResultTB::bindingEnergy(int stateindex, double gap) const {
    if (stateindex < 0 || stateindex >= static_cast<int>(eigval.size())) {
        throw std::invalid_argument("bindingEnergy: state index out of range");
    }
    if (gap == -1) {
        gap = determineGap();
    } else if (gap < 0) {
        std::cout << "Provided gap value must be positive" << std::endl;
    }
    return eigval[stateindex] - gap;
}

// Result.tpp:106-139: Re sum_k |A_k|^2 e^{-i k.(r_e - r_h)}, |A_k|^2 the weight of k.
double ResultTB::ftExcitonEnvelope(int stateindex, const std::vector<double>& re, const std::vector<double>& rh) const {
    const std::vector<cdouble> c = state_column(*this, stateindex, "ftExcitonEnvelope");
    if (re.size() != 3 || rh.size() != 3) throw std::invalid_argument("ftExcitonEnvelope: positions are 3-vectors");
    const SystemTB& s = *exciton->system;
    const long long blk = nrows / s.nk;
    cdouble ft(0.0, 0.0);
    for (int k = 0; k < s.nk; ++k) {
        double w = 0.0;
        for (long long b = 0; b < blk; ++b) w += std::norm(c[k * blk + b]);
        const double* kp = s.kpoints.data() + 3 * static_cast<size_t>(k);
        const double ph = kp[0] * (re[0] - rh[0]) + kp[1] * (re[1] - rh[1]) + kp[2] * (re[2] - rh[2]);
        ft += w * cdouble(std::cos(ph), -std::sin(ph));
    }
    return ft.real();
}

// Result.tpp:492-499: the larger of the largest kx and ky of the mesh.
double ResultTB::boundingBoxBZ() const {
    const SystemTB& s = *exciton->system;
    double mx = -INFINITY, my = -INFINITY;
    for (int k = 0; k < s.nk; ++k) {
        mx = std::max(mx, s.kpoints[3 * static_cast<size_t>(k)]);
        my = std::max(my, s.kpoints[3 * static_cast<size_t>(k) + 1]);
    }
    return (mx > my) ? mx : my;
}

namespace {

// Centred combinations of the ndim reciprocal vectors with coefficients -1, 0, 1 (first
// coefficient fastest), as Lattice::generateCombinations(3, ndim, true).
std::vector<std::vector<double>> reciprocal_shifts(const SystemTB& s) {
    const std::vector<double> comb = s.generateCombinations(3, s.ndim, true);
    std::vector<std::vector<double>> out;
    for (size_t i = 0; i < comb.size() / s.ndim; ++i) {
        std::vector<double> g(3, 0.0);
        for (int j = 0; j < s.ndim; ++j)
            for (int x = 0; x < 3; ++x) g[x] += comb[i * s.ndim + j] * s.reciprocalLattice[3 * j + x];
        out.push_back(g);
    }
    return out;
}

double mesh_step(const SystemTB& s) {
    if (s.nk < 2) throw std::invalid_argument("needs at least two k-points");
    double d = 0.0;
    for (int x = 0; x < 3; ++x) d += (s.kpoints[3 + x] - s.kpoints[x]) * (s.kpoints[3 + x] - s.kpoints[x]);
    return std::sqrt(d);
}

}  // namespace

// Result.tpp:199-225: the k-space amplitude repeated over the neighbouring reciprocal
// cells, kept where |kx| and |ky| are below boundingBoxBZ().
void ResultTB::writeExtendedReciprocalAmplitude(const std::vector<cdouble>& coefs, FILE* f) const {
    check_state(*this, coefs, "writeExtendedReciprocalAmplitude");
    const SystemTB& s = *exciton->system;
    const long long nbc = static_cast<long long>(exciton->valenceBands.size() * exciton->conductionBands.size());
    const double box = boundingBoxBZ();
    const double dk = mesh_step(s);
    const std::vector<std::vector<double>> shifts = reciprocal_shifts(s);
    for (int k = 0; k < s.nk; ++k) {
        double a = 0.0;
        for (long long b = 0; b < nbc; ++b) a += std::abs(coefs[nbc * k + b]) * std::abs(coefs[nbc * k + b]);
        a /= dk;
        const double* kp = s.kpoints.data() + 3 * static_cast<size_t>(k);
        for (const auto& g : shifts) {
            const double x = kp[0] + g[0], y = kp[1] + g[1], z = kp[2] + g[2];
            if (std::abs(x) < box && std::abs(y) < box) fprintf(f, "%11.8lf\t%11.8lf\t%11.8lf\t%11.8lf\n", x, y, z, a);
        }
    }
    fprintf(f, "#\n");
}

void ResultTB::writeExtendedReciprocalAmplitude(int stateindex, FILE* f) const {
    writeExtendedReciprocalAmplitude(state_column(*this, stateindex, "writeExtendedReciprocalAmplitude"), f);
}

// Result.tpp:249-268 (one valence and one conduction band only)
void ResultTB::writePhase(const std::vector<cdouble>& coefs, FILE* f) const {
    if (exciton->bandList.size() != 2) throw std::logic_error("writePhase requires only one valence and conduction bands");
    check_state(*this, coefs, "writePhase");
    const SystemTB& s = *exciton->system;
    fprintf(f, "kx\tky\tkz\tMod.\tArg.\n");
    for (int k = 0; k < s.nk; ++k) {
        const double* kp = s.kpoints.data() + 3 * static_cast<size_t>(k);
        fprintf(f, "%11.8lf\t%11.8lf\t%11.8lf\t%11.8lf\t%11.8lf\n", kp[0], kp[1], kp[2], std::abs(coefs[k]),
                std::arg(coefs[k]));
    }
    fprintf(f, "#\n");
}

void ResultTB::writePhase(int stateindex, FILE* f) const { writePhase(state_column(*this, stateindex, "writePhase"), f); }

// Result.tpp:288-321: modulus / |k_1 - k_0| and phase over the neighbouring reciprocal
// cells (the shifts use the first two reciprocal vectors, as the reference does).
void ResultTB::writeExtendedPhase(const std::vector<cdouble>& coefs, FILE* f) const {
    if (exciton->bandList.size() != 2) {
        throw std::logic_error("writeExtendedPhase requires only one valence and conduction bands");
    }
    check_state(*this, coefs, "writeExtendedPhase");
    const SystemTB& s = *exciton->system;
    if (s.ndim < 2) throw std::logic_error("writeExtendedPhase: needs a two-dimensional reciprocal lattice");
    fprintf(f, "kx\tky\tkz\tMod.\tArg.\n");
    const double box = boundingBoxBZ();
    const double dk = mesh_step(s);
    const std::vector<double> comb = s.generateCombinations(3, s.ndim, true);
    for (int k = 0; k < s.nk; ++k) {
        const double mod = std::abs(coefs[k]) / dk;
        const double ph = std::arg(coefs[k]);
        const double* kp = s.kpoints.data() + 3 * static_cast<size_t>(k);
        for (size_t i = 0; i < comb.size() / s.ndim; ++i) {
            double g[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < 2; ++j)
                for (int x = 0; x < 3; ++x) g[x] += comb[i * s.ndim + j] * s.reciprocalLattice[3 * j + x];
            const double x = kp[0] + g[0], y = kp[1] + g[1], z = kp[2] + g[2];
            if (std::abs(x) < box && std::abs(y) < box) {
                fprintf(f, "%11.8lf\t%11.8lf\t%11.8lf\t%11.8lf\t%11.8lf\n", x, y, z, mod, ph);
            }
        }
    }
    fprintf(f, "#\n");
}

void ResultTB::writeExtendedPhase(int stateindex, FILE* f) const {
    writeExtendedPhase(state_column(*this, stateindex, "writeExtendedPhase"), f);
}

}  // namespace xatu
