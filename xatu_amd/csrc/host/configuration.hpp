// Parsers for Xatu's '#'-block text formats: system (.model) files and
// exciton configuration files. Behaviour follows the reference parsers:
//   ConfigurationBase     src/ConfigurationBase.cpp:24-198
//   SystemConfiguration   src/SystemConfiguration.cpp:20-258
//   ExcitonConfiguration  src/ExcitonConfiguration.cpp:21-156
// Storage is plain std::vector (no Armadillo); matrices are column-major like
// Armadillo's, so the buffers can be handed to the C-ABI unchanged.
#pragma once

#include <complex>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace xatu {

using cdouble = std::complex<double>;

class ConfigurationBase {
public:
    std::string filename;
    std::vector<std::string> foundArguments;
    std::vector<std::string> expectedArguments;
    std::map<std::string, std::vector<std::string>> contents;

    virtual ~ConfigurationBase() = default;

protected:
    explicit ConfigurationBase(const std::string& file);
    // the file's lines, without their '\n'
    const std::vector<std::string>& fileLines() const { return lines_; }

    void extractArguments();
    void extractRawContent();
    void checkArguments() const;

    static std::string parseArgument(const std::string& line);
    static std::string parseWord(const std::string& content);
    static std::string standarizeLine(std::string line);

    template <typename T>
    static T parseScalar(const std::string& line) {
        T value{};
        std::istringstream iss(line);
        iss >> value;
        return value;
    }

    template <typename T>
    static std::vector<T> parseLine(const std::string& line) {
        std::vector<T> values;
        std::istringstream iss(line);
        T value;
        while (iss >> value) values.push_back(value);
        return values;
    }

private:
    std::vector<std::string> lines_;  // file split into lines
    bool endsWithNewline_ = true;     // whether the last line had a trailing '\n'
};

// Parsed contents of a system (.model) file (SystemConfiguration.hpp:13-32).
struct SystemInfo {
    int ndim = 0;
    double filling = 0.0;
    std::vector<double> bravaisLattice;   // ndim x 3, row-major (row = lattice vector)
    std::vector<double> motif;            // natoms x 4, row-major (x, y, z, species)
    std::vector<int> norbitals;           // orbitals per species
    std::vector<double> bravaisVectors;   // ncells x 3, row-major
    int matrixDim = 0;                    // dimension of each H(R) block
    std::vector<cdouble> hamiltonian;     // matrixDim x matrixDim x ncells, column-major slices
    std::vector<cdouble> overlap;         // same layout, empty if absent
    int nHamiltonian = 0;
    int nOverlap = 0;
    int natoms() const { return static_cast<int>(motif.size() / 4); }
    int ncells() const { return static_cast<int>(bravaisVectors.size() / 3); }
};

class SystemConfiguration : public ConfigurationBase {
public:
    SystemInfo systemInfo;
    explicit SystemConfiguration(const std::string& filename);

protected:
    // Reads the file without parsing it as a '#'-block model (for other formats).
    struct RawFile {};
    SystemConfiguration(const std::string& filename, RawFile) : ConfigurationBase(filename) {}

private:
    void parseContent();
    void checkContentCoherence() const;
    static std::vector<double> parseVectors(const std::vector<std::string>& content);
    static std::vector<double> parseMotif(const std::vector<std::string>& content);
    static std::vector<int> parseOrbitals(const std::vector<std::string>& content);
    static std::vector<cdouble> parseMatrices(const std::vector<std::string>& content, int& dim, int& count);
};

// Parsed contents of an exciton file (ExcitonConfiguration.hpp:18-40).
struct ExcitonInfo {
    std::string label;
    int ncell = 0;
    int nbands = 0;
    int submeshFactor = 1;
    std::vector<long long> bands;
    std::vector<double> Q{0.0, 0.0, 0.0};
    std::vector<double> shift;
    double cutoff = 0.0;
    std::vector<double> eps;
    std::string mode = "realspace";
    bool exchange = false;
    double scissor = 0.0;
    int nReciprocalVectors = 0;
    std::string potential = "keldysh";
    std::string exchangePotential = "keldysh";
    double regularization = 0.0;
};

class ExcitonConfiguration : public ConfigurationBase {
public:
    ExcitonInfo excitonInfo;
    std::vector<std::string> supportedPotentials{"keldysh", "coulomb"};
    explicit ExcitonConfiguration(const std::string& filename);

private:
    void parseContent();
    void checkContentCoherence() const;
};

}  // namespace xatu
