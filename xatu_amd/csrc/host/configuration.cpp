#include "configuration.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <iostream>
#include <stdexcept>

namespace xatu {

// ConfigurationBase(file): reference src/ConfigurationBase.cpp:24-33.
ConfigurationBase::ConfigurationBase(const std::string& file) : filename(file) {
    if (file.empty()) {
        throw std::invalid_argument("ConfigurationBase: file must not be empty");
    }
    std::ifstream in(file.c_str(), std::ios::binary);
    if (!in.is_open()) {
        throw std::invalid_argument("ConfigurationBase: file does not exist");
    }
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    lines_.clear();
    size_t start = 0;
    while (start <= text.size()) {
        size_t pos = text.find('\n', start);
        if (pos == std::string::npos) {
            lines_.push_back(text.substr(start));
            break;
        }
        lines_.push_back(text.substr(start, pos - start));
        start = pos + 1;
    }
    // A trailing '\n' leaves one empty string that getline() would never return.
    endsWithNewline_ = !text.empty() && text.back() == '\n';
    if (endsWithNewline_ || text.empty()) lines_.pop_back();
}

// parseArgument: text after the first '#', whitespace removed, lower-cased
// (src/ConfigurationBase.cpp:43-49).
std::string ConfigurationBase::parseArgument(const std::string& line) {
    std::size_t pos = line.find('#');
    std::string str = line.substr(pos + 1);
    str.erase(std::remove_if(str.begin(), str.end(), [](unsigned char ch) { return std::isspace(ch); }),
              str.end());
    std::transform(str.begin(), str.end(), str.begin(), [](unsigned char ch) { return std::tolower(ch); });
    return str;
}

std::string ConfigurationBase::parseWord(const std::string& content) {
    return parseArgument("#" + content);
}

// ',' and ';' are value separators like ' ' (src/ConfigurationBase.cpp:139-147).
std::string ConfigurationBase::standarizeLine(std::string line) {
    std::replace(line.begin(), line.end(), ',', ' ');
    std::replace(line.begin(), line.end(), ';', ' ');
    return line;
}

// Every line holding '#' names an argument, except the file's last line
// (the reference tests peek() != EOF after reading it, :57-68).
void ConfigurationBase::extractArguments() {
    foundArguments.clear();
    for (size_t i = 0; i < lines_.size(); ++i) {
        const bool atEOF = (i + 1 == lines_.size());
        if (lines_[i].find('#') != std::string::npos && !atEOF) {
            foundArguments.push_back(parseArgument(lines_[i]));
        }
    }
}

// Block contents keyed by argument (src/ConfigurationBase.cpp:92-119). Empty
// lines and lines holding '!' are skipped; a block is stored when the next
// '#' line starts, or at end of file when the last line has no newline.
void ConfigurationBase::extractRawContent() {
    std::string arg;
    std::vector<std::string> content;
    for (size_t i = 0; i < lines_.size(); ++i) {
        const std::string& line = lines_[i];
        if (line.find('#') != std::string::npos) {
            if (!content.empty()) {
                contents[arg] = content;
                content.clear();
            }
            arg = parseArgument(line);
        } else if (line.empty() || line.find('!') != std::string::npos) {
            // skipped
        } else {
            content.push_back(standarizeLine(line));
        }
        const bool eofSet = (i + 1 == lines_.size()) && !endsWithNewline_;
        if (eofSet) contents[arg] = content;
    }
}

void ConfigurationBase::checkArguments() const {
    if (expectedArguments.empty()) {
        throw std::logic_error("Expected arguments must be defined first");
    }
    for (const auto& arg : expectedArguments) {
        if (std::find(foundArguments.begin(), foundArguments.end(), arg) == foundArguments.end()) {
            throw std::logic_error("Missing arguments in config. file: missing " + arg);
        }
    }
}

// ---------------------------------------------------------------------------
// SystemConfiguration (src/SystemConfiguration.cpp:20-258)

SystemConfiguration::SystemConfiguration(const std::string& filename) : ConfigurationBase(filename) {
    expectedArguments = {"dimension", "bravaislattice", "motif", "norbitals",
                         "filling", "bravaisvectors", "hamiltonian"};
    parseContent();
    checkArguments();
    checkContentCoherence();
}

void SystemConfiguration::parseContent() {
    extractArguments();
    extractRawContent();
    if (contents.empty()) {
        throw std::logic_error("File contents must be extracted first");
    }
    for (const auto& arg : foundArguments) {
        std::vector<std::string> content = contents[arg];
        if (arg == "dimension") {
            if (content.size() != 1) throw std::logic_error("Expected only one line for 'Dimension'");
            systemInfo.ndim = parseScalar<int>(content[0]);
        } else if (arg == "bravaislattice") {
            systemInfo.bravaisLattice = parseVectors(content);
        } else if (arg == "motif") {
            systemInfo.motif = parseMotif(content);
        } else if (arg == "norbitals") {
            systemInfo.norbitals = parseOrbitals(content);
        } else if (arg == "filling") {
            if (content.size() != 1) throw std::logic_error("Expected only one line in 'filling' field");
            systemInfo.filling = parseScalar<double>(content[0]);
            double intpart;
            if (std::modf(systemInfo.filling, &intpart) != 0) {
                throw std::invalid_argument("Filling must be a positive integer.");
            }
        } else if (arg == "bravaisvectors") {
            systemInfo.bravaisVectors = parseVectors(content);
        } else if (arg == "hamiltonian") {
            systemInfo.hamiltonian = parseMatrices(content, systemInfo.matrixDim, systemInfo.nHamiltonian);
        } else if (arg == "overlap") {
            int dim = 0;
            systemInfo.overlap = parseMatrices(content, dim, systemInfo.nOverlap);
        } else {
            std::cout << "Unexpected argument: " << arg << ", skipping block..." << std::endl;
        }
    }
}

std::vector<double> SystemConfiguration::parseVectors(const std::vector<std::string>& content) {
    std::vector<double> out;
    for (const auto& line : content) {
        std::vector<double> v = parseLine<double>(line);
        if (v.size() != 3) throw std::logic_error("Bravais vectors must have three components");
        out.insert(out.end(), v.begin(), v.end());
    }
    return out;
}

std::vector<double> SystemConfiguration::parseMotif(const std::vector<std::string>& content) {
    std::vector<double> out;
    for (const auto& line : content) {
        std::istringstream iss(line);
        double x, y, z, species;
        if ((iss >> x >> y >> z >> species).fail()) {
            throw std::invalid_argument("Motif must be of shape (x,y,z,'species')");
        }
        out.insert(out.end(), {x, y, z, species});
    }
    return out;
}

std::vector<int> SystemConfiguration::parseOrbitals(const std::vector<std::string>& content) {
    if (content.size() != 1) throw std::invalid_argument("Error: Orbital information must be one line only");
    return parseLine<int>(content[0]);
}

// Dense real or complex matrices separated by '&' lines; complex values are
// written as "re im" token pairs (src/SystemConfiguration.cpp:159-228). The
// first line decides real vs complex; the lines before the first '&' give
// the dimension.
std::vector<cdouble> SystemConfiguration::parseMatrices(const std::vector<std::string>& content, int& dim,
                                                        int& count) {
    if (content.empty()) throw std::logic_error("Matrix block is empty");
    const std::string& first = content[0];
    const bool isComplex = first.find('i') != std::string::npos || first.find('j') != std::string::npos;
    int ndim = 0;
    for (const auto& line : content) {
        if (line.find('&') != std::string::npos) break;
        ndim++;
    }
    std::vector<cdouble> out;
    std::vector<cdouble> matrix(static_cast<size_t>(ndim) * ndim, cdouble(0.0, 0.0));
    int row = 0;
    count = 0;
    for (const auto& line : content) {
        if (line.find('&') != std::string::npos) {
            out.insert(out.end(), matrix.begin(), matrix.end());
            std::fill(matrix.begin(), matrix.end(), cdouble(0.0, 0.0));
            row = 0;
            count++;
            continue;
        }
        if (row >= ndim) throw std::logic_error("Matrix has more rows than columns");
        std::istringstream iss(line);
        std::string tok;
        int j = 0;
        std::vector<cdouble> rowv(ndim, cdouble(0.0, 0.0));
        while (iss >> tok) {
            double re = 0.0, im = 0.0;
            {
                std::istringstream vs(tok);
                vs >> re;
            }
            if (isComplex) {
                iss >> tok;
                std::istringstream vs(tok);
                vs >> im;
            }
            if (j >= ndim) throw std::logic_error("Matrix row has more entries than the matrix dimension");
            rowv[j++] = cdouble(re, im);
        }
        for (int c = 0; c < ndim; ++c) matrix[row + static_cast<size_t>(c) * ndim] = rowv[c];
        row++;
    }
    dim = ndim;
    return out;
}

void SystemConfiguration::checkContentCoherence() const {
    const int nlat = static_cast<int>(systemInfo.bravaisLattice.size() / 3);
    if (systemInfo.ndim != nlat) {
        throw std::invalid_argument("Error: Dimensions must match number of Bravais basis");
    }
    if (systemInfo.nHamiltonian != systemInfo.ncells()) {
        throw std::invalid_argument("Error: Number of H matrices must match number of Bravais vectors");
    }
    if (!systemInfo.overlap.empty() && systemInfo.nOverlap != systemInfo.nHamiltonian) {
        throw std::invalid_argument("Error: Number of overlap matrices must match number of H matrices");
    }
    std::vector<int> species;
    for (int a = 0; a < systemInfo.natoms(); ++a) {
        int s = static_cast<int>(systemInfo.motif[4 * a + 3]);
        if (std::find(species.begin(), species.end(), s) == species.end()) species.push_back(s);
    }
    if (systemInfo.norbitals.size() != species.size()) {
        throw std::invalid_argument(
            "Error: Number of different species must match be consistent in motif and orbitals");
    }
}

// ---------------------------------------------------------------------------
// ExcitonConfiguration (src/ExcitonConfiguration.cpp:21-156)

ExcitonConfiguration::ExcitonConfiguration(const std::string& filename) : ConfigurationBase(filename) {
    expectedArguments = {"ncells", "dielectric"};
    parseContent();
    checkArguments();
    checkContentCoherence();
}

void ExcitonConfiguration::parseContent() {
    extractArguments();
    extractRawContent();
    if (contents.empty()) throw std::logic_error("File contents must be extracted first");
    for (const auto& arg : foundArguments) {
        std::vector<std::string> content = contents[arg];
        if (content.empty()) continue;
        if (content.size() != 1) throw std::logic_error("Expected only one line per field");
        const std::string& line = content[0];
        if (arg == "label") {
            excitonInfo.label = standarizeLine(line);
        } else if (arg == "ncells") {
            excitonInfo.ncell = parseScalar<int>(line);
        } else if (arg == "submesh") {
            excitonInfo.submeshFactor = parseScalar<int>(line);
        } else if (arg == "shift") {
            excitonInfo.shift = parseLine<double>(line);
        } else if (arg == "bands") {
            excitonInfo.nbands = parseScalar<int>(line);
        } else if (arg == "bandlist") {
            excitonInfo.bands = parseLine<long long>(line);
        } else if (arg == "totalmomentum") {
            excitonInfo.Q = parseLine<double>(line);
        } else if (arg == "cutoff") {
            excitonInfo.cutoff = parseScalar<double>(line);
        } else if (arg == "dielectric") {
            excitonInfo.eps = parseLine<double>(line);
        } else if (arg == "reciprocal") {
            excitonInfo.mode = "reciprocalspace";
            excitonInfo.nReciprocalVectors = parseScalar<int>(line);
        } else if (arg == "exchange") {
            std::string s = parseWord(line);
            if (s != "true" && s != "false") {
                throw std::invalid_argument("Exchange option must be set to 'true' or 'false'.");
            }
            if (s == "true") excitonInfo.exchange = true;
        } else if (arg == "exchange.potential") {
            excitonInfo.exchangePotential = parseWord(line);
        } else if (arg == "potential") {
            excitonInfo.potential = parseWord(line);
        } else if (arg == "scissor") {
            excitonInfo.scissor = parseScalar<double>(line);
        } else if (arg == "regularization") {
            excitonInfo.regularization = parseScalar<double>(line);
        } else {
            std::cout << "Unexpected argument: " << arg << ", skipping block..." << std::endl;
        }
    }
}

void ExcitonConfiguration::checkContentCoherence() const {
    const ExcitonInfo& e = excitonInfo;
    if (e.Q.size() != 3) throw std::logic_error("'Q' must be a 3d vector");
    if (e.ncell <= 0) throw std::logic_error("'ncell' must be a positive number");
    if (e.bands.empty() && e.nbands == 0) throw std::logic_error("'bands' must be specified");
    if (e.eps.empty()) throw std::logic_error("'dielectric' must be specified");
    bool potentialFound = false, exchangeFound = false;
    for (const auto& p : supportedPotentials) {
        if (e.potential == p) potentialFound = true;
        if (e.exchange && e.exchangePotential == p) exchangeFound = true;
    }
    if (!potentialFound) throw std::invalid_argument("Specified 'potential' not supported. Use 'keldysh' or 'coulomb'");
    if (e.exchange && !exchangeFound) {
        throw std::invalid_argument("Specified 'exchange.potential' not supported. Use 'keldysh' or 'coulomb'");
    }
    if (e.mode != "realspace" && e.mode != "reciprocalspace") {
        throw std::invalid_argument("Invalid mode. Use 'realspace' or 'reciprocalspace'");
    }
}

}  // namespace xatu
