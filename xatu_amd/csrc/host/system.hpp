// Lattice, k-mesh and tight-binding system (host side, FP64).
// Mirrors the reference classes Lattice (src/Lattice.cpp), System (src/System.cpp)
// and SystemTB (src/SystemTB.cpp) without Armadillo: vectors are std::vector with
// rows of 3 doubles (row-major), complex matrices column-major.
#pragma once

#include <array>
#include <complex>
#include <vector>

#include "configuration.hpp"

namespace xatu {

constexpr double kPI = 3.141592653589793;

class Lattice {
public:
    int ndim = 0;
    int natoms = 0;
    int ncells = 0;
    int nk = 0;
    int factor = 1;
    double a = 0.0;             // |a1| (Lattice.cpp:97)
    double c = 1.0;             // layer height (Lattice.cpp:99-110)
    double unitCellArea = 0.0;  // Lattice.cpp:119-131
    std::vector<double> bravaisLattice;     // ndim x 3
    std::vector<double> reciprocalLattice;  // ndim x 3
    std::vector<double> motif;              // natoms x 4
    std::vector<double> unitCellList;       // ncells x 3
    std::vector<double> kpoints;            // nk x 3
    std::vector<double> meshBZ;             // nmesh x 3
    // Integer offset of every k-point from k-point 0 along each reciprocal vector,
    // in units of 1/(n*factor) (n*factor = points per axis of meshBZ).
    std::vector<int> kpointOffsets;         // nk x ndim
    int meshPoints = 0;                     // points per axis of meshBZ

    void brillouinZoneMesh(int n);
    void reducedBrillouinZoneMesh(int n, int factor);
    void shiftBZ(const std::vector<double>& shift);
    void calculateInverseReciprocalMatrix();
    int findEquivalentPointBZ(const double* kpoint, int ncell) const;
    std::vector<double> generateCombinations(int nvalues, int dim, bool centered = false) const;
    std::vector<double> truncateSupercell(int ncell, double radius) const;

protected:
    void initializeLatticeAttributes(const SystemConfiguration& configuration);
    std::vector<double> inverseReciprocalMatrix;  // ndim x ndim
};

class SystemTB : public Lattice {
public:
    explicit SystemTB(const SystemConfiguration& configuration);

    int basisdim = 0;
    int filling = 0;
    int fermiLevel = 0;                 // filling - 1 (System.cpp:52)
    std::vector<int> orbitals;          // per species
    std::vector<cdouble> hamiltonianMatrices;  // basisdim x basisdim x ncells
    std::vector<cdouble> overlapMatrices;      // same, empty if absent
    bool isTriangular = false;
    bool isAU = false;

    void setFilling(int filling);
    void setTriangular(bool t) { isTriangular = t; }
    void setAU(bool au) { isAU = au; }

    // orbitals of each atom of the motif, in basis order
    std::vector<int> orbitalsPerAtom() const;
    // H(k) = sum_R H(R) exp(i k.R) (SystemTB.cpp:60-75)
    std::vector<cdouble> hamiltonian(const double* k) const;
    std::vector<cdouble> overlap(const double* k) const;
    // Eigenvalues ascending and eigenvectors of H(k), Lowdin-orthogonalised when an
    // overlap is present (SystemTB.cpp:113-156).
    void solveBands(const double* k, std::vector<double>& eigval, std::vector<cdouble>& eigvec) const;
};

}  // namespace xatu
