// Reader of CRYSTAL (.outp) SCF outputs, the input that CRYSTALConfiguration
// (src/CRYSTALConfiguration.cpp) turns into a tight-binding system.
//
// What the reader relies on in an output file:
//  * header lines "KEY   value ..." for the atom, shell, AO, electron and core-electron
//    counts (the first integer after the key);
//  * "DIRECT LATTICE VECTOR COMPONENTS (ANGSTROM)" followed by the three lattice vectors;
//    a non-periodic direction is printed as (0, 0, 500) (then (0, 500, 0)) or as zero, so
//    such rows lower the dimension; the last such block in the file wins;
//  * the atom table (its header names both ATOM and SHELL): a rule of asterisks, then one
//    row per atom "index  Z-code  symbol  shells  x  y  z ..."; species are numbered by the
//    first appearance of their symbol;
//  * the basis-set section ("LOCAL ATOMIC FUNCTIONS BASIS SET", then three header lines):
//    one row per atom "index symbol x y z"; the first atom of a species is followed by its
//    shells, each a header "n TYPE" or "first-  last TYPE" (n, last: global index of the
//    shell's last atomic orbital) and rows of four numbers (exponent and the s, p, d/f/g
//    coefficients), which CRYSTAL may print without a blank between them
//    ("2.926E+00-5.119E-01"); later atoms of a species list no shells;
//  * matrices "OVERLAP MATRIX - CELL N.  k( i j l)" and "FOCK MATRIX - CELL N. ...": the
//    lower triangle in blocks of columns, each block a blank line, a line of column indices
//    and rows "row  value value ..."; the matrix ends with its (n, n) entry. Cell k is kept
//    if k <= ncells; its lattice vector is i a1 + j a2 + l a3 over the periodic directions;
//  * unrestricted (magnetic) runs print the alpha Fock matrices, then a line naming BETA
//    ELECTRONS, then the beta ones; spin-orbit runs are recognised by a marker and their
//    Fock matrices are not read (as the reference does).
// Without a lattice block the system is finite, which is refused.
#include "crystal.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace xatu {

namespace {

constexpr const char* kLatticeKey = "DIRECT LATTICE VECTOR COMPONENTS (ANGSTROM)";
constexpr const char* kBasisKey = "LOCAL ATOMIC FUNCTIONS BASIS SET";
constexpr const char* kOverlapKey = "OVERLAP MATRIX - CELL N.";
constexpr const char* kFockKey = "FOCK MATRIX - CELL N.";
constexpr const char* kSpinOrbitKey = "to_be_defined_for_crystal23";
constexpr const char* kUnrestrictedKey = "UNRESTRICTED OPEN SHELL";

bool has(const std::string& s, const char* key) { return s.find(key) != std::string::npos; }

// Numbers at the start of s, left to right; strtod also separates numbers printed
// without a blank between them. Stops at the first text that is not a number.
std::vector<double> leading_numbers(const std::string& s) {
    std::vector<double> out;
    const char* p = s.c_str();
    while (true) {
        char* end = nullptr;
        const double v = std::strtod(p, &end);
        if (end == p) break;
        out.push_back(v);
        p = end;
    }
    return out;
}

// The first integer after key on the line (0 if there is none).
int int_after(const std::string& line, const char* key) {
    const size_t at = line.find(key);
    if (at == std::string::npos) return 0;
    return static_cast<int>(std::strtol(line.c_str() + at + std::char_traits<char>::length(key), nullptr, 10));
}

// Whitespace-separated words of a line.
std::vector<std::string> words(const std::string& s) {
    std::vector<std::string> w;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
        const size_t b = i;
        while (i < s.size() && !std::isspace(static_cast<unsigned char>(s[i]))) ++i;
        if (i > b) w.push_back(s.substr(b, i - b));
    }
    return w;
}

// Cursor over the lines of the file.
class Cursor {
public:
    explicit Cursor(const std::vector<std::string>& lines) : lines_(lines) {}
    bool more() const { return at_ < lines_.size(); }
    const std::string& take() {
        static const std::string empty;
        return more() ? lines_[at_++] : empty;
    }
    void skip(int n) {
        for (int i = 0; i < n; ++i) take();
    }
    size_t where() const { return at_; }
    void back_to(size_t w) { at_ = w; }

private:
    const std::vector<std::string>& lines_;
    size_t at_ = 0;
};

// Cell tag of a matrix title: "... CELL N.   2( -1 -1  0)" -> 2 and (-1, -1, 0).
struct CellTag {
    int index = 0;
    int m[3] = {0, 0, 0};
};

CellTag cell_tag(const std::string& title, const char* key) {
    CellTag t;
    const char* p = title.c_str() + title.find(key) + std::char_traits<char>::length(key);
    char* end = nullptr;
    t.index = static_cast<int>(std::strtol(p, &end, 10));
    p = end;
    while (*p && *p != '(') ++p;
    if (*p == '(') ++p;
    for (int& c : t.m) {
        c = static_cast<int>(std::strtol(p, &end, 10));
        p = end;
    }
    return t;
}

// Lower triangle of an n x n matrix printed in column blocks; returns the full matrix
// column-major with the upper entries zero.
std::vector<cdouble> read_lower_triangle(Cursor& cur, int n) {
    std::vector<cdouble> m(static_cast<size_t>(n) * n, cdouble(0.0, 0.0));
    std::vector<int> cols;
    while (cur.more()) {
        const std::string& line = cur.take();
        if (line.empty()) {  // a new column block: its column indices come next
            cols.clear();
            for (double c : leading_numbers(cur.take())) cols.push_back(static_cast<int>(c));
            continue;
        }
        const std::vector<double> v = leading_numbers(line);
        if (v.empty()) continue;
        const int row = static_cast<int>(v[0]);
        if (v.size() - 1 > cols.size()) throw std::logic_error("CRYSTAL matrix row longer than its column header");
        int last_col = 0;
        for (size_t i = 1; i < v.size(); ++i) {
            last_col = cols[i - 1];
            if (row < 1 || row > n || last_col < 1 || last_col > n) {
                throw std::logic_error("CRYSTAL matrix index outside the number of AO");
            }
            m[(row - 1) + static_cast<size_t>(last_col - 1) * n] = v[i];
        }
        if (row == n && last_col == n) break;
    }
    return m;
}

// |r - (x, y, z)| within 0.1 per component.
bool near(const double* r, double x, double y, double z) {
    return std::fabs(r[0] - x) <= 0.1 && std::fabs(r[1] - y) <= 0.1 && std::fabs(r[2] - z) <= 0.1;
}

// Periodic dimension of the three printed lattice vectors (rows of v).
int periodic_dimension(const std::vector<double>& v) {
    if (!near(&v[6], 0, 0, 500) && !near(&v[6], 0, 0, 0)) return 3;
    if (!near(&v[3], 0, 500, 0) && !near(&v[3], 0, 0, 0)) return 2;
    return 1;
}

// Interleave alpha and beta spin blocks: out = kron(A, diag(1, 0)) + kron(B, diag(0, 1)),
// n x n blocks to 2n x 2n, column-major.
std::vector<cdouble> spin_blocks(const cdouble* A, const cdouble* B, int n) {
    const int n2 = 2 * n;
    std::vector<cdouble> out(static_cast<size_t>(n2) * n2, cdouble(0.0, 0.0));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            out[(2 * i) + static_cast<size_t>(2 * j) * n2] = A[i + static_cast<size_t>(j) * n];
            out[(2 * i + 1) + static_cast<size_t>(2 * j + 1) * n2] = B[i + static_cast<size_t>(j) * n];
        }
    return out;
}

}  // namespace

// State of one pass over the file.
struct CRYSTALConfiguration::Scan {
    CRYSTALConfiguration& out;
    Cursor cur;
    int ncells;
    bool have_lattice = false;
    bool beta = false;  // past the line that starts the beta-spin matrices
    std::vector<std::string> species;  // symbols in order of first appearance

    // atom table: a rule, then one row per atom
    void atoms() {
        if (out.natoms == 0) throw std::logic_error("Must parse first number of atoms");
        cur.skip(1);
        species.clear();
        out.motif.assign(static_cast<size_t>(out.natoms) * 4, 0.0);
        out.shellsPerSpecies.clear();
        out.atomic_number_ordering.clear();
        for (int a = 0; a < out.natoms; ++a) {
            const std::vector<std::string> w = words(cur.take());
            if (w.size() < 7) throw std::logic_error("CRYSTAL atom table: short row");
            const std::string& symbol = w[2];
            int sp = 0;
            while (sp < static_cast<int>(species.size()) && species[sp] != symbol) ++sp;
            if (sp == static_cast<int>(species.size())) {
                species.push_back(symbol);
                out.shellsPerSpecies.push_back(std::atoi(w[3].c_str()));
                out.atomic_number_ordering.push_back(std::atoi(w[1].c_str()));
            }
            for (int x = 0; x < 3; ++x) out.motif[4 * a + x] = std::strtod(w[4 + x].c_str(), nullptr);
            out.motif[4 * a + 3] = sp;
        }
        out.nspecies = static_cast<int>(species.size());
    }

    // basis-set section: three header lines, then the atoms with their species' shells
    void basis() {
        cur.skip(3);
        std::vector<bool> seen(out.shellsPerSpecies.size(), false);
        out.orbitalsPerSpecies.assign(out.shellsPerSpecies.size(), 0);
        int ao_before = 0;  // atomic orbitals of the atoms before the current one
        for (int a = 0; a < out.natoms; ++a) {
            cur.take();  // "index symbol x y z"
            const int sp = static_cast<int>(out.motif[4 * a + 3]);
            if (seen.at(sp)) {
                ao_before += out.orbitalsPerSpecies[sp];
                continue;
            }
            seen[sp] = true;
            std::vector<std::vector<std::vector<double>>> shells;
            int last_ao = ao_before;
            for (int sh = 0; sh < out.shellsPerSpecies.at(sp); ++sh) {
                // "n TYPE" or "first-  last TYPE": the last integer is the shell's last AO
                for (const std::string& w : words(cur.take())) {
                    char* end = nullptr;
                    const long v = std::strtol(w.c_str(), &end, 10);
                    if (end != w.c_str()) last_ao = static_cast<int>(v);
                }
                std::vector<std::vector<double>> rows;
                while (cur.more()) {
                    const size_t mark = cur.where();
                    std::vector<double> v = leading_numbers(cur.take());
                    if (v.size() != 4) {
                        cur.back_to(mark);
                        break;
                    }
                    rows.push_back(std::move(v));
                }
                shells.push_back(std::move(rows));
            }
            out.gaussianCoefficients[sp] = std::move(shells);
            out.orbitalsPerSpecies[sp] = last_ao - ao_before;
            ao_before = last_ao;
        }
    }

    void lattice() {
        std::vector<double> v(9, 0.0);
        for (int r = 0; r < 3; ++r) {
            const std::vector<double> x = leading_numbers(cur.take());
            for (int c = 0; c < 3 && c < static_cast<int>(x.size()); ++c) v[3 * r + c] = x[c];
        }
        out.ndim = periodic_dimension(v);
        v.resize(3 * static_cast<size_t>(out.ndim));
        out.bravaisLattice = v;
        have_lattice = true;
    }

    void overlap(const std::string& title) {
        const CellTag t = cell_tag(title, kOverlapKey);
        if (t.index > ncells) return;
        double R[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < out.ndim; ++i)
            for (int x = 0; x < 3; ++x) R[x] += t.m[i] * out.bravaisLattice[3 * i + x];
        out.bravaisVectors.insert(out.bravaisVectors.end(), R, R + 3);
        const std::vector<cdouble> m = read_lower_triangle(cur, out.norbitals);
        out.overlapMatrices.insert(out.overlapMatrices.end(), m.begin(), m.end());
        ++out.nOverlap;
    }

    void fock(const std::string& title) {
        if (cell_tag(title, kFockKey).index > ncells) return;
        const std::vector<cdouble> m = read_lower_triangle(cur, out.norbitals);
        std::vector<cdouble>* dst = nullptr;
        if (out.MAGNETIC_FLAG) {
            dst = beta ? &out.betaMatrices : &out.alphaMatrices;
            ++(beta ? out.nBeta : out.nAlpha);
        } else if (!out.SOC_FLAG) {
            dst = &out.fockMatrices;
            ++out.nFock;
        }
        if (dst) dst->insert(dst->end(), m.begin(), m.end());
    }

    void run() {
        while (cur.more()) {
            const std::string& line = cur.take();
            if (has(line, kLatticeKey)) {
                lattice();
            } else if (has(line, "N. OF ATOMS PER CELL")) {
                out.natoms = int_after(line, "N. OF ATOMS PER CELL");
            } else if (has(line, "NUMBER OF SHELLS")) {
                out.nsh = int_after(line, "NUMBER OF SHELLS");
            } else if (has(line, "NUMBER OF AO")) {
                out.norbitals = int_after(line, "NUMBER OF AO");
            } else if (has(line, "N. OF ELECTRONS PER CELL")) {
                out.totalElectrons = int_after(line, "N. OF ELECTRONS PER CELL");
            } else if (has(line, "CORE ELECTRONS PER CELL")) {
                out.coreElectrons = int_after(line, "CORE ELECTRONS PER CELL");
            } else if (has(line, "ATOM") && has(line, "SHELL")) {
                atoms();
            } else if (has(line, kBasisKey)) {
                basis();
            } else if (has(line, kOverlapKey)) {
                overlap(line);
            } else if (has(line, kFockKey)) {
                fock(line);
            } else if (has(line, kSpinOrbitKey)) {
                out.SOC_FLAG = true;
            } else if (has(line, kUnrestrictedKey)) {
                out.MAGNETIC_FLAG = true;
            }
            if (has(line, "BETA") && has(line, "ELECTRONS")) {
                beta = true;
                out.alpha_electrons = false;
            }
        }
        if (!have_lattice) {
            out.ndim = 0;
            throw std::logic_error("Finite systems are not yet implemented");
        }
    }
};

CRYSTALConfiguration::CRYSTALConfiguration(const std::string& file, int ncells)
    : SystemConfiguration(file, RawFile{}) {
    Scan scan{*this, Cursor(fileLines()), ncells};
    scan.run();
    fillSystemInfo();
}

// The system the tight-binding code takes: the matrices are in Hartree (SystemTB::setAU)
// and printed as lower triangles (SystemTB detects that); a magnetic run becomes one
// spinful system with alpha and beta spin blocks and the overlap repeated for both spins.
void CRYSTALConfiguration::fillSystemInfo() {
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
    const bool spinful = SOC_FLAG || MAGNETIC_FLAG;
    if (spinful) {
        s.filling *= 2;
        for (int& n : s.norbitals) n *= 2;
    }
    if (MAGNETIC_FLAG) {
        if (nAlpha != nBeta || nAlpha != nOverlap) {
            throw std::logic_error("CRYSTAL output: alpha, beta and overlap matrices differ in number");
        }
        const size_t nn = static_cast<size_t>(norbitals) * norbitals;
        std::vector<cdouble> F, S;
        for (int c = 0; c < nAlpha; ++c) {
            const std::vector<cdouble> f = spin_blocks(&alphaMatrices[c * nn], &betaMatrices[c * nn], norbitals);
            const std::vector<cdouble> o = spin_blocks(&overlapMatrices[c * nn], &overlapMatrices[c * nn], norbitals);
            F.insert(F.end(), f.begin(), f.end());
            S.insert(S.end(), o.begin(), o.end());
        }
        fockMatrices = std::move(F);
        nFock = nAlpha;
        s.overlap = std::move(S);
        s.matrixDim = 2 * norbitals;
    }
    s.hamiltonian = fockMatrices;
    s.nHamiltonian = nFock;
}

}  // namespace xatu
