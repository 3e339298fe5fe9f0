// Reader of CRYSTAL (.outp) SCF outputs: lattice, motif, basis sizes, and the overlap
// and Fock (Kohn-Sham) matrices of the first ncells cells, the data members
// CRYSTALConfiguration (include/xatu/CRYSTALConfiguration.hpp) exposes. The result
// is a SystemConfiguration whose systemInfo feeds SystemTB (non-orthogonal basis: the
// bands are Lowdin-orthogonalised; the matrices are printed as lower triangles and are
// in Hartree, see SystemTB::setAU).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "configuration.hpp"

namespace xatu {

class CRYSTALConfiguration : public SystemConfiguration {
public:
    explicit CRYSTALConfiguration(const std::string& file, int ncells = 20);

    int ndim = 0, natoms = 0, norbitals = 0, nsh = 0, nspecies = 0;
    int totalElectrons = 0, coreElectrons = 0;
    bool SOC_FLAG = false;
    bool MAGNETIC_FLAG = false;
    bool alpha_electrons = true;
    std::vector<double> motif;            // natoms x 4 (x, y, z, species)
    std::vector<double> bravaisLattice;   // ndim x 3
    std::vector<int> atomic_number_ordering, shellsPerSpecies, orbitalsPerSpecies;
    // per species: per shell, rows {exponent, s, p, d/f/g coefficient}
    std::map<int, std::vector<std::vector<std::vector<double>>>> gaussianCoefficients;
    std::vector<double> bravaisVectors;   // cells x 3
    std::vector<cdouble> overlapMatrices, fockMatrices, alphaMatrices, betaMatrices;  // norbitals^2 per cell
    int nOverlap = 0, nFock = 0, nAlpha = 0, nBeta = 0;

private:
    struct Scan;  // one pass over the file's lines
    void fillSystemInfo();
};

}  // namespace xatu
