// CRYSTAL (.outp) reader (reference: src/CRYSTALConfiguration.cpp:18-464). The file is
// scanned line by line for the markers CRYSTAL prints; each block is parsed the way the
// reference parses it (stream extraction of the same fields), so quirks of the format
// (shell ranges written "2-     4 P", exponents and coefficients without a separating
// blank) are read identically.
#include "crystal.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>
#include <stdexcept>

namespace xatu {

namespace {

const char* kSocString = "to_be_defined_for_crystal23";
const char* kMagneticString = "UNRESTRICTED OPEN SHELL";

bool contains(const std::string& line, const char* key) { return line.find(key) != std::string::npos; }

// The text after the first occurrence of key (as std::string::substr with the
// reference's position arithmetic).
std::string after(const std::string& line, const char* key) {
    const size_t pos = line.find(key);
    const size_t start = pos + std::char_traits<char>::length(key);
    return start <= line.size() ? line.substr(start) : std::string();
}

template <typename T>
T scalar_after(const std::string& line, const char* key) {
    std::istringstream iss(after(line, key));
    T v{};
    iss >> v;
    return v;
}

// Number of leading numbers a line holds (src/CRYSTALConfiguration.cpp:455-464).
size_t count_numbers(const std::string& txt) {
    std::istringstream iss(txt);
    double token;
    size_t n = 0;
    while (iss >> token) ++n;
    return n;
}

bool approx_row(const double* r, double x, double y, double z) {
    return std::fabs(r[0] - x) <= 0.1 && std::fabs(r[1] - y) <= 0.1 && std::fabs(r[2] - z) <= 0.1;
}

}  // namespace

CRYSTALConfiguration::CRYSTALConfiguration(const std::string& file, int ncells)
    : SystemConfiguration(file, RawFile{}) {
    parseContent(ncells);
    mapContent();
}

bool CRYSTALConfiguration::getline(std::string& line) {
    const std::vector<std::string>& lines = fileLines();
    if (pos_ >= lines.size()) return false;
    line = lines[pos_++];
    return true;
}

// src/CRYSTALConfiguration.cpp:31-170
void CRYSTALConfiguration::parseContent(int ncells) {
    int countr = 0;
    std::string line;
    while (getline(line)) {
        if (contains(line, "DIRECT LATTICE VECTOR COMPONENTS (ANGSTROM)")) {
            parseBravaisLattice();
            countr++;
        } else if (contains(line, "N. OF ATOMS PER CELL")) {
            natoms = scalar_after<int>(line, "N. OF ATOMS PER CELL");
        } else if (contains(line, "NUMBER OF SHELLS")) {
            nsh = scalar_after<int>(line, "NUMBER OF SHELLS");
        } else if (contains(line, "NUMBER OF AO")) {
            norbitals = scalar_after<int>(line, "NUMBER OF AO");
        } else if (contains(line, "N. OF ELECTRONS PER CELL")) {
            totalElectrons = scalar_after<int>(line, "N. OF ELECTRONS PER CELL");
        } else if (contains(line, "CORE ELECTRONS PER CELL")) {
            coreElectrons = scalar_after<int>(line, "CORE ELECTRONS PER CELL");
        } else if (contains(line, "ATOM") && contains(line, "SHELL")) {
            if (natoms == 0) throw std::logic_error("Must parse first number of atoms");
            parseAtoms();
        } else if (contains(line, "LOCAL ATOMIC FUNCTIONS BASIS SET")) {
            parseAtomicBasis();
        } else if (contains(line, "OVERLAP MATRIX")) {
            std::istringstream iss(after(line, "OVERLAP MATRIX - CELL N."));
            int cellIndex = 0, x = 0, y = 0, z = 0;
            std::string parenthesis;
            iss >> cellIndex >> parenthesis >> x >> y >> z;
            const int coef[3] = {x, y, z};
            if (cellIndex <= ncells) {
                double cell[3] = {0.0, 0.0, 0.0};
                for (int i = 0; i < ndim; ++i)
                    for (int j = 0; j < 3; ++j) cell[j] += bravaisLattice[3 * i + j] * coef[i];
                bravaisVectors.insert(bravaisVectors.end(), cell, cell + 3);
                const std::vector<cdouble> m = parseMatrix();
                overlapMatrices.insert(overlapMatrices.end(), m.begin(), m.end());
                nOverlap++;
            }
        } else if (contains(line, kSocString)) {
            SOC_FLAG = true;
        } else if (contains(line, kMagneticString)) {
            MAGNETIC_FLAG = true;
        }

        if (contains(line, "BETA") && contains(line, "ELECTRONS")) alpha_electrons = false;

        if (contains(line, "FOCK MATRIX")) {
            std::istringstream iss(after(line, "FOCK MATRIX - CELL N."));
            int cellIndex = 0, x = 0, y = 0, z = 0;
            std::string parenthesis;
            iss >> cellIndex >> parenthesis >> x >> y >> z;
            if (cellIndex <= ncells) {
                const std::vector<cdouble> m = parseMatrix();
                if (MAGNETIC_FLAG) {
                    if (alpha_electrons) {
                        alphaMatrices.insert(alphaMatrices.end(), m.begin(), m.end());
                        nAlpha++;
                    } else {
                        betaMatrices.insert(betaMatrices.end(), m.begin(), m.end());
                        nBeta++;
                    }
                } else if (SOC_FLAG) {
                    continue;
                } else {
                    fockMatrices.insert(fockMatrices.end(), m.begin(), m.end());
                    nFock++;
                }
            }
        }
    }
    if (countr == 0) {
        ndim = 0;
        throw std::logic_error("Finite systems are not yet implemented");
    }
}

// src/CRYSTALConfiguration.cpp:179-188
void CRYSTALConfiguration::parseBravaisLattice() {
    std::vector<double> v;
    std::string line;
    for (int i = 0; i < 3; ++i) {
        getline(line);
        std::istringstream iss(standarizeLine(line));
        double a = 0.0, b = 0.0, c = 0.0;
        iss >> a >> b >> c;
        v.insert(v.end(), {a, b, c});
    }
    bravaisLattice = v;
    extractDimension();
}

// CRYSTAL writes a non-periodic direction as (0, 0, 500) (then (0, 500, 0)); such a
// vector, or a zero one, lowers the dimension (src/CRYSTALConfiguration.cpp:194-210).
void CRYSTALConfiguration::extractDimension() {
    int rows = 3;
    const double* r2 = bravaisLattice.data() + 6;
    if (approx_row(r2, 0.0, 0.0, 500.0) || approx_row(r2, 0.0, 0.0, 0.0)) {
        rows = 2;
        const double* r1 = bravaisLattice.data() + 3;
        if (approx_row(r1, 0.0, 500.0, 0.0) || approx_row(r1, 0.0, 0.0, 0.0)) rows = 1;
    }
    bravaisLattice.resize(3 * static_cast<size_t>(rows));
    ndim = rows;
}

// src/CRYSTALConfiguration.cpp:214-247: per atom "index Z species shells x y z"; species
// are numbered in order of first appearance.
void CRYSTALConfiguration::parseAtoms() {
    std::string line;
    std::vector<std::string> species;
    std::vector<int> shells, atomic;
    std::vector<double> m(static_cast<size_t>(natoms) * 4, 0.0);
    getline(line);  // asterisks
    for (int i = 0; i < natoms; ++i) {
        getline(line);
        std::istringstream iss(line);
        int index = 0, atomic_number = 0, ns = 0;
        double x = 0.0, y = 0.0, z = 0.0;
        std::string chemical;
        iss >> index >> atomic_number >> chemical >> ns >> x >> y >> z;
        auto it = std::find(species.begin(), species.end(), chemical);
        int sidx;
        if (it == species.end()) {
            sidx = static_cast<int>(species.size());
            species.push_back(chemical);
            shells.push_back(ns);
            atomic.push_back(atomic_number);
        } else {
            sidx = static_cast<int>(it - species.begin());
        }
        m[4 * i + 0] = x;
        m[4 * i + 1] = y;
        m[4 * i + 2] = z;
        m[4 * i + 3] = sidx;
    }
    motif = m;
    shellsPerSpecies = shells;
    atomic_number_ordering = atomic;
    nspecies = static_cast<int>(species.size());
}

// src/CRYSTALConfiguration.cpp:255-327: for each new species, its shells ("n TYPE" or
// "first- last TYPE") and their gaussian rows of four numbers; the orbitals of a species
// are the last AO index of its shells minus the AO count of the atoms before it.
void CRYSTALConfiguration::parseAtomicBasis() {
    std::string line;
    for (int i = 0; i < 3; ++i) getline(line);  // asterisks, header, asterisks
    std::vector<std::string> species;
    int norb = 0, totalOrbitals = 0, ns = 0;
    for (int atomIndex = 0; atomIndex < natoms; ++atomIndex) {
        getline(line);
        std::istringstream iss(line);
        int natom = 0;
        std::string chemical;
        iss >> natom >> chemical;
        if (std::find(species.begin(), species.end(), chemical) == species.end()) {
            species.push_back(chemical);
        } else {
            totalOrbitals += orbitalsPerSpecies.at(static_cast<int>(motif[4 * atomIndex + 3]));
            continue;
        }
        std::vector<std::vector<std::vector<double>>> shellsCoefs;
        for (int shell = 0; shell < shellsPerSpecies.at(ns); ++shell) {
            getline(line);
            std::istringstream s2(line);
            std::string shellType;
            s2 >> norb >> shellType;
            if (shellType == "-") s2 >> norb >> shellType;
            std::vector<std::vector<double>> coefs;
            size_t previous = pos_;
            while (getline(line)) {
                if (count_numbers(line) != 4) {
                    pos_ = previous;
                    break;
                }
                std::istringstream s3(line);
                double e = 0.0, sc = 0.0, pc = 0.0, dc = 0.0;
                s3 >> e >> sc >> pc >> dc;
                coefs.push_back({e, sc, pc, dc});
                previous = pos_;
            }
            shellsCoefs.push_back(coefs);
            gaussianCoefficients[ns] = shellsCoefs;
        }
        orbitalsPerSpecies.push_back(norb - totalOrbitals);
        totalOrbitals = norb;
        ns++;
    }
}

// src/CRYSTALConfiguration.cpp:334-370: blocks of columns (a blank line, the column
// indices, the rows "i  v_i,c1 v_i,c2 ..."), until the entry (norbitals, norbitals).
std::vector<cdouble> CRYSTALConfiguration::parseMatrix() {
    std::vector<cdouble> m(static_cast<size_t>(norbitals) * norbitals, cdouble(0.0, 0.0));
    std::vector<int> cols;
    std::string line;
    int rowIndex = 0, colIndex = 0;
    while (getline(line)) {
        if (line.empty()) {
            cols.clear();
            std::string idx;
            getline(idx);
            std::istringstream iss(idx);
            int c;
            while (iss >> c) cols.push_back(c);
            getline(line);
            if (line.empty()) continue;
        }
        std::istringstream iss(line);
        iss >> rowIndex;
        double coef;
        size_t i = 0;
        while (iss >> coef) {
            if (i >= cols.size()) throw std::logic_error("CRYSTAL matrix row longer than its column header");
            colIndex = cols[i];
            if (rowIndex < 1 || rowIndex > norbitals || colIndex < 1 || colIndex > norbitals) {
                throw std::logic_error("CRYSTAL matrix index outside the number of AO");
            }
            m[(rowIndex - 1) + static_cast<size_t>(colIndex - 1) * norbitals] = coef;
            ++i;
        }
        if (rowIndex == norbitals && colIndex == norbitals) return m;
    }
    return m;
}

// src/CRYSTALConfiguration.cpp:377-414
void CRYSTALConfiguration::mapContent() {
    SystemInfo& s = systemInfo;
    s.ndim = ndim;
    s.bravaisLattice = bravaisLattice;
    s.motif = motif;
    s.filling = totalElectrons / 2;
    s.bravaisVectors = bravaisVectors;
    s.overlap = overlapMatrices;
    s.nOverlap = nOverlap;
    s.norbitals = orbitalsPerSpecies;
    s.matrixDim = norbitals;
    if (SOC_FLAG) {
        s.filling *= 2;
        for (int& n : s.norbitals) n *= 2;
    } else if (MAGNETIC_FLAG) {
        s.filling *= 2;
        for (int& n : s.norbitals) n *= 2;
        if (nAlpha != nBeta || nAlpha != nOverlap) {
            throw std::logic_error("CRYSTAL output: alpha, beta and overlap matrices differ in number");
        }
        const int n = norbitals, n2 = 2 * norbitals;
        const size_t nn = static_cast<size_t>(n) * n, nn2 = static_cast<size_t>(n2) * n2;
        fockMatrices.assign(nn2 * nAlpha, cdouble(0.0, 0.0));
        std::vector<cdouble> ov(nn2 * nAlpha, cdouble(0.0, 0.0));
        for (int c = 0; c < nAlpha; ++c)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    // kron(alpha, diag(1, 0)) + kron(beta, diag(0, 1)); kron(S, 1_2)
                    fockMatrices[c * nn2 + (2 * i) + static_cast<size_t>(2 * j) * n2] = alphaMatrices[c * nn + i + j * n];
                    fockMatrices[c * nn2 + (2 * i + 1) + static_cast<size_t>(2 * j + 1) * n2] = betaMatrices[c * nn + i + j * n];
                    ov[c * nn2 + (2 * i) + static_cast<size_t>(2 * j) * n2] = overlapMatrices[c * nn + i + j * n];
                    ov[c * nn2 + (2 * i + 1) + static_cast<size_t>(2 * j + 1) * n2] = overlapMatrices[c * nn + i + j * n];
                }
        nFock = nAlpha;
        s.overlap = ov;
        s.matrixDim = n2;
    }
    s.hamiltonian = fockMatrices;
    s.nHamiltonian = nFock;
}

}  // namespace xatu
