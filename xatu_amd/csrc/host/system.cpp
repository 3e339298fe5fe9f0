#include "system.hpp"

#include <cmath>
#include <iostream>
#include <stdexcept>

#include "../common/heev.hpp"

namespace xatu {

namespace {

// Solve the n x n system A x = rhs (row-major A), n <= 3, by Gaussian elimination
// with partial pivoting.
std::vector<double> solve_small(std::vector<double> A, std::vector<double> rhs, int n) {
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A[r * n + col]) > std::fabs(A[piv * n + col])) piv = r;
        if (A[piv * n + col] == 0.0) throw std::runtime_error("Failed to obtain reciprocal lattice vectors");
        if (piv != col) {
            for (int j = 0; j < n; ++j) std::swap(A[col * n + j], A[piv * n + j]);
            std::swap(rhs[col], rhs[piv]);
        }
        for (int r = col + 1; r < n; ++r) {
            const double f = A[r * n + col] / A[col * n + col];
            for (int j = col; j < n; ++j) A[r * n + j] -= f * A[col * n + j];
            rhs[r] -= f * rhs[col];
        }
    }
    std::vector<double> x(n);
    for (int r = n - 1; r >= 0; --r) {
        double s = rhs[r];
        for (int j = r + 1; j < n; ++j) s -= A[r * n + j] * x[j];
        x[r] = s / A[r * n + r];
    }
    return x;
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

bool check_if_triangular(const std::vector<cdouble>& H, int n) {
    // is_trimatu() != is_trimatl() (src/utils.cpp:176-178)
    bool upper = true, lower = true;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const cdouble v = H[i + static_cast<size_t>(j) * n];
            if (i > j && v != cdouble(0.0, 0.0)) upper = false;
            if (i < j && v != cdouble(0.0, 0.0)) lower = false;
        }
    return upper != lower;
}

}  // namespace

void Lattice::initializeLatticeAttributes(const SystemConfiguration& configuration) {
    const SystemInfo& s = configuration.systemInfo;
    ndim = s.ndim;
    bravaisLattice = s.bravaisLattice;
    motif = s.motif;
    unitCellList = s.bravaisVectors;
    natoms = s.natoms();
    ncells = s.ncells();
    nk = 0;

    // reciprocal lattice (Lattice.cpp:56-80)
    reciprocalLattice.assign(static_cast<size_t>(ndim) * 3, 0.0);
    if (ndim == 1) {
        const double n2 = bravaisLattice[0] * bravaisLattice[0] + bravaisLattice[1] * bravaisLattice[1] +
                          bravaisLattice[2] * bravaisLattice[2];
        for (int j = 0; j < 3; ++j) reciprocalLattice[j] = 2. * kPI * bravaisLattice[j] / n2;
    } else {
        std::vector<double> A(static_cast<size_t>(ndim) * ndim);
        for (int r = 0; r < ndim; ++r)
            for (int cc = 0; cc < ndim; ++cc) A[r * ndim + cc] = bravaisLattice[3 * r + cc];
        for (int i = 0; i < ndim; ++i) {
            std::vector<double> rhs(ndim, 0.0);
            rhs[i] = 2 * kPI;
            std::vector<double> x = solve_small(A, rhs, ndim);
            for (int j = 0; j < ndim; ++j) reciprocalLattice[3 * i + j] = x[j];
        }
    }
    // lattice parameters (Lattice.cpp:87-111)
    a = norm3(bravaisLattice.data());
    const double ref = motif.empty() ? 0.0 : motif[2];
    double cz = 0.0;
    for (int i = 0; i < natoms; ++i) cz = std::max(cz, std::fabs(motif[4 * i + 2] - ref));
    c = (cz == 0) ? 1 : cz;
    // unit cell area (Lattice.cpp:119-131)
    if (ndim == 1) {
        unitCellArea = a;
    } else {
        const double* a1 = bravaisLattice.data();
        const double* a2 = bravaisLattice.data() + 3;
        const double cr[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2],
                              a1[0] * a2[1] - a1[1] * a2[0]};
        if (ndim == 2) {
            unitCellArea = norm3(cr);
        } else {
            const double* a3 = bravaisLattice.data() + 6;
            unitCellArea = cr[0] * a3[0] + cr[1] * a3[1] + cr[2] * a3[2];
        }
    }
}

std::vector<double> Lattice::generateCombinations(int nvalues, int dim, bool centered) const {
    long long total = 1;
    for (int j = 0; j < dim; ++j) total *= nvalues;
    const int shift = centered ? nvalues / 2 : 0;
    std::vector<double> out(static_cast<size_t>(total) * dim);
    for (long long i = 0; i < total; ++i) {
        long long rem = i;
        for (int j = 0; j < dim; ++j) {
            out[static_cast<size_t>(i) * dim + j] = static_cast<double>(rem % nvalues) - shift;
            rem /= nvalues;
        }
    }
    return out;
}

// Monkhorst-Pack mesh centred at Gamma (Lattice.cpp:140-162).
void Lattice::brillouinZoneMesh(int n) {
    if (n < 1) throw std::invalid_argument("brillouinZoneMesh: n must be positive");
    std::vector<double> comb = generateCombinations(n, ndim);
    const long long total = static_cast<long long>(comb.size() / ndim);
    kpoints.assign(static_cast<size_t>(total) * 3, 0.0);
    kpointOffsets.assign(comb.size(), 0);
    for (long long i = 0; i < total; ++i) {
        for (int j = 0; j < ndim; ++j) {
            double cj = comb[i * ndim + j];
            kpointOffsets[i * ndim + j] = static_cast<int>(cj);
            if (n % 2 == 1) cj += 1. / 2;
            const double f = (2 * cj - n) / (2 * n);
            for (int x = 0; x < 3; ++x) kpoints[3 * i + x] += f * reciprocalLattice[3 * j + x];
        }
    }
    meshBZ = kpoints;
    nk = static_cast<int>(total);
    meshPoints = n;
}

// Sub-mesh of n points per axis of the mesh with n*factor points (Lattice.cpp:172-197).
void Lattice::reducedBrillouinZoneMesh(int n, int fac) {
    if (fac < 1) throw std::invalid_argument("reducedBrillouinZoneMesh: factor must be positive");
    brillouinZoneMesh(n * fac);
    std::vector<double> comb = generateCombinations(n, ndim);
    const long long total = static_cast<long long>(comb.size() / ndim);
    kpoints.assign(static_cast<size_t>(total) * 3, 0.0);
    kpointOffsets.assign(comb.size(), 0);
    for (long long i = 0; i < total; ++i) {
        for (int j = 0; j < ndim; ++j) {
            double cj = comb[i * ndim + j];
            kpointOffsets[i * ndim + j] = static_cast<int>(cj);
            if (n % 2 == 1) cj += 1. / 2;
            const double f = (2 * cj - n) / (2 * n * fac);
            for (int x = 0; x < 3; ++x) kpoints[3 * i + x] += f * reciprocalLattice[3 * j + x];
        }
    }
    nk = static_cast<int>(total);
    factor = fac;
    meshPoints = n * fac;
}

void Lattice::shiftBZ(const std::vector<double>& shift) {
    if (shift.size() != 3) {
        std::cout << "shift vector must be 3d" << std::endl;
        return;
    }
    if (kpoints.empty()) {
        std::cout << "To call this method kpoints must be initiallized first" << std::endl;
        return;
    }
    for (int i = 0; i < nk; ++i)
        for (int x = 0; x < 3; ++x) kpoints[3 * i + x] += shift[x];
}

void Lattice::calculateInverseReciprocalMatrix() {
    std::vector<double> G(static_cast<size_t>(ndim) * ndim);
    for (int i = 0; i < ndim; ++i)
        for (int j = 0; j < ndim; ++j) {
            double s = 0;
            for (int x = 0; x < 3; ++x) s += reciprocalLattice[3 * i + x] * reciprocalLattice[3 * j + x];
            G[i * ndim + j] = s;
        }
    inverseReciprocalMatrix.assign(static_cast<size_t>(ndim) * ndim, 0.0);
    for (int col = 0; col < ndim; ++col) {
        std::vector<double> e(ndim, 0.0);
        e[col] = 1.0;
        std::vector<double> x = solve_small(G, e, ndim);
        for (int r = 0; r < ndim; ++r) inverseReciprocalMatrix[r * ndim + col] = x[r];
    }
}

// Index in meshBZ of the point equivalent to kpoint (Lattice.cpp:251-279).
int Lattice::findEquivalentPointBZ(const double* kpoint, int ncell) const {
    if (inverseReciprocalMatrix.empty()) {
        throw std::logic_error("findEquivalentPointBZ: call calculateInverseReciprocalMatrix first");
    }
    ncell = ncell * factor;
    std::vector<double> indep(ndim), coefs(ndim);
    for (int i = 0; i < ndim; ++i) {
        double s = 0;
        for (int x = 0; x < 3; ++x) s += reciprocalLattice[3 * i + x] * kpoint[x];
        indep[i] = s;
    }
    for (int i = 0; i < ndim; ++i) {
        double s = 0;
        for (int j = 0; j < ndim; ++j) s += inverseReciprocalMatrix[i * ndim + j] * indep[j];
        coefs[i] = s * 2 * ncell;
    }
    int index = 0;
    const int cells_array[3] = {1, ncell, ncell * ncell};
    for (int i = 0; i < ndim; ++i) {
        double cf = (ncell % 2 == 1) ? coefs[i] - 1 : coefs[i];
        cf = std::round(cf);
        if (cf >= ncell) {
            cf -= 2 * ncell;
        } else if (cf < -ncell) {
            cf += 2 * ncell;
        }
        cf += ncell;
        cf /= 2;
        index += static_cast<int>(cf * cells_array[i]);
    }
    return index;
}

// Cells of the centred supercell inside a sphere (Lattice.cpp:320-341).
std::vector<double> Lattice::truncateSupercell(int ncell, double radius) const {
    std::vector<double> comb = generateCombinations(ncell, ndim, true);
    const long long total = static_cast<long long>(comb.size() / ndim);
    std::vector<double> out;
    for (long long i = 0; i < total; ++i) {
        double v[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < ndim; ++j)
            for (int x = 0; x < 3; ++x) v[x] += comb[i * ndim + j] * bravaisLattice[3 * j + x];
        if (norm3(v) < radius + 1E-5) out.insert(out.end(), v, v + 3);
    }
    return out;
}

// ---------------------------------------------------------------------------

SystemTB::SystemTB(const SystemConfiguration& configuration) {
    initializeLatticeAttributes(configuration);
    const SystemInfo& s = configuration.systemInfo;
    orbitals = s.norbitals;
    hamiltonianMatrices = s.hamiltonian;
    overlapMatrices = s.overlap;
    filling = static_cast<int>(s.filling);
    fermiLevel = filling - 1;
    basisdim = 0;
    for (int i = 0; i < natoms; ++i) {
        const int species = static_cast<int>(motif[4 * i + 3]);
        if (species < 0 || species >= static_cast<int>(orbitals.size())) {
            throw std::invalid_argument("motif species index has no orbital count");
        }
        basisdim += orbitals[species];
    }
    if (s.matrixDim != basisdim) {
        throw std::invalid_argument("Error: Hamiltonian dimension does not match the number of orbitals");
    }
    isTriangular = ncells > 0 && check_if_triangular(hamiltonianMatrices, basisdim);
}

void SystemTB::setFilling(int f) {
    if (f > 0) {
        filling = f;
        fermiLevel = f - 1;
    } else {
        std::cout << "Filling must be a positive integer" << std::endl;
    }
}

std::vector<int> SystemTB::orbitalsPerAtom() const {
    std::vector<int> out(natoms);
    for (int i = 0; i < natoms; ++i) out[i] = orbitals[static_cast<int>(motif[4 * i + 3])];
    return out;
}

static std::vector<cdouble> bloch_sum(const std::vector<cdouble>& mats, const std::vector<double>& cells, int n,
                                      int ncells, bool triangular, const double* k) {
    std::vector<cdouble> h(static_cast<size_t>(n) * n, cdouble(0.0, 0.0));
    const size_t nn = static_cast<size_t>(n) * n;
    for (int r = 0; r < ncells; ++r) {
        const double* cell = cells.data() + 3 * r;
        const double ph = k[0] * cell[0] + k[1] * cell[1] + k[2] * cell[2];
        const cdouble e = std::exp(cdouble(0.0, ph));
        const cdouble* m = mats.data() + nn * r;
        for (size_t i = 0; i < nn; ++i) h[i] += m[i] * e;
    }
    if (triangular) {
        for (int i = 0; i < n; ++i) h[i + static_cast<size_t>(i) * n] -= h[i + static_cast<size_t>(i) * n] / 2.0;
        std::vector<cdouble> ht(h);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) h[i + static_cast<size_t>(j) * n] += std::conj(ht[j + static_cast<size_t>(i) * n]);
    }
    return h;
}

std::vector<cdouble> SystemTB::hamiltonian(const double* k) const {
    return bloch_sum(hamiltonianMatrices, unitCellList, basisdim, ncells, isTriangular, k);
}

std::vector<cdouble> SystemTB::overlap(const double* k) const {
    return bloch_sum(overlapMatrices, unitCellList, basisdim, ncells, isTriangular, k);
}

void SystemTB::solveBands(const double* k, std::vector<double>& eigval, std::vector<cdouble>& eigvec) const {
    const int n = basisdim;
    std::vector<cdouble> h = hamiltonian(k);
    if (!overlapMatrices.empty()) {
        if (isAU) {
            for (auto& v : h) v *= 27.2;
        }
        // Lowdin orthogonalisation: h <- S^{-1/2} h S^{-1/2} (SystemTB.cpp:136-156)
        std::vector<cdouble> s = overlap(k);
        std::vector<double> sw(n);
        std::vector<cdouble> sv(static_cast<size_t>(n) * n);
        xatu_linalg::heev(n, s.data(), n, sw.data(), sv.data());
        for (double& x : sw) {
            if (!(x > 0.0)) throw std::invalid_argument("Zero or negative overlap eigenvalues found, exiting...");
            x = 1. / std::sqrt(x);
        }
        std::vector<cdouble> root(static_cast<size_t>(n) * n, cdouble(0.0, 0.0));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                cdouble acc(0.0, 0.0);
                for (int l = 0; l < n; ++l)
                    acc += sv[i + static_cast<size_t>(l) * n] * sw[l] * std::conj(sv[j + static_cast<size_t>(l) * n]);
                root[i + static_cast<size_t>(j) * n] = acc;
            }
        std::vector<cdouble> t(static_cast<size_t>(n) * n, cdouble(0.0, 0.0));
        for (int j = 0; j < n; ++j)
            for (int l = 0; l < n; ++l) {
                const cdouble hl = h[l + static_cast<size_t>(j) * n];
                for (int i = 0; i < n; ++i) t[i + static_cast<size_t>(j) * n] += root[i + static_cast<size_t>(l) * n] * hl;
            }
        std::fill(h.begin(), h.end(), cdouble(0.0, 0.0));
        for (int j = 0; j < n; ++j)
            for (int l = 0; l < n; ++l) {
                const cdouble rl = root[l + static_cast<size_t>(j) * n];
                for (int i = 0; i < n; ++i) h[i + static_cast<size_t>(j) * n] += t[i + static_cast<size_t>(l) * n] * rl;
            }
    }
    eigval.resize(n);
    eigvec.resize(static_cast<size_t>(n) * n);
    xatu_linalg::heev(n, h.data(), n, eigval.data(), eigvec.data());
}

}  // namespace xatu
