// ExcitonTB / ResultTB facade: the reference's public API
// (include/xatu/ExcitonTB.hpp:24-157, include/xatu/Exciton.hpp:40-105,
// include/xatu/Result.hpp:11-66) backed by the C-ABI of include/xatu_bse.h.
// The single-particle problem (bands and phase-fixed eigenvectors per k) runs on
// the host; the lattice Fourier sums, the BSE assembly and the diagonalisation run
// on the GPU with the matrix resident in HBM.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/xatu_bse.h"
#include "system.hpp"

namespace xatu {

class ResultTB;

class ExcitonTB {
public:
    // Constructors (ExcitonTB.hpp:72-88, src/ExcitonTB.cpp:91-210).
    ExcitonTB(const SystemConfiguration& config, int ncell = 20, int nbands = 1, int nrmbands = 0,
              const std::vector<double>& parameters = {1, 5, 1}, const std::vector<double>& Q = {0., 0., 0.});
    ExcitonTB(const SystemConfiguration& config, int ncell, const std::vector<long long>& bands,
              const std::vector<double>& parameters = {1, 5, 1}, const std::vector<double>& Q = {0., 0., 0.});
    ExcitonTB(const SystemConfiguration& config, const ExcitonConfiguration& excitonConfig);
    ExcitonTB(std::shared_ptr<SystemTB> sys, int ncell = 20, int nbands = 1, int nrmbands = 0,
              const std::vector<double>& parameters = {1, 5, 1}, const std::vector<double>& Q = {0., 0., 0.});
    ExcitonTB(std::shared_ptr<SystemTB> sys, int ncell, const std::vector<long long>& bands,
              const std::vector<double>& parameters = {1, 5, 1}, const std::vector<double>& Q = {0., 0., 0.});
    ~ExcitonTB();
    ExcitonTB(const ExcitonTB&) = delete;
    ExcitonTB& operator=(const ExcitonTB&) = delete;

    // Setters (ExcitonTB.hpp:93-98, Exciton.hpp:97-103)
    void setParameters(const std::vector<double>& parameters);
    void setParameters(double eps_m, double eps_s, double r0);
    void setGauge(const std::string& gauge);
    void setMode(const std::string& mode);
    void setReciprocalVectors(int n);
    void setRegularization(double regularization);
    void setUnitCells(int ncell);
    void setBands(const std::vector<long long>& bands);
    void setBands(int nbands, int nrmbands);
    void setQ(const std::vector<double>& Q);
    void setCutoff(double cutoff);
    void setScissor(double shift);
    void setExchange(bool exchange);
    void setPotential(const std::string& potential);
    void setExchangePotential(const std::string& potential);
    // Build options of this implementation (no counterpart in the reference):
    // "fp64" (default) or "fp32" arithmetic for the assembly and the products;
    // matrix-free products; Davidson block size; HIP device.
    void setPrecision(const std::string& precision);
    void setMatrixFree(bool matrixFree);
    void setDavidsonBlock(int block);
    void setDavidsonTolerance(double tol);
    void setDevice(int device);

    // Pipeline (ExcitonTB.hpp:140-143)
    void brillouinZoneMesh(int ncell);
    void initializeHamiltonian();
    void BShamiltonian();
    void BShamiltonian(const std::vector<long long>& basis);  // nbasis x 3 row-major {v, c, k}
    std::unique_ptr<ResultTB> diagonalize(const std::string& method = "diag", int nstates = 8);

    // Read-only state (Exciton.hpp:52-86)
    std::shared_ptr<SystemTB> system;
    int ncell = 0, nbands = 0, nrmbands = 0;
    long long totalCells = 0;
    long long excitonbasisdim = 0;
    double eps_m = 1, eps_s = 5, r0 = 1, regularization = 0, cutoff = 0, scissor = 0;
    bool exchange = false;
    std::string gauge = "lattice", mode = "realspace", potential = "keldysh", exchangePotential = "keldysh";
    int nReciprocalVectors = 1;
    std::vector<long long> bands, valenceBands, conductionBands;
    std::vector<long long> bandList;
    std::vector<double> Q{0., 0., 0.};
    std::vector<long long> basisStates;        // N x 3 row-major {v, c, k}
    std::vector<double> eigvalKStack, eigvalKQStack;    // nb x nk
    std::vector<cdouble> eigvecKStack, eigvecKQStack;   // basisdim x nb x nk
    std::map<int, int> bandToIndex;
    std::string precision = "fp64";
    bool matrixFree = false;
    int davidsonBlock = 0;
    double davidsonTol = 0.0;
    int device = 0;

    // The resident BSE matrix copied to the host as a full column-major matrix of the
    // current basis (hbsDim() x hbsDim(): the full basis, or the subset given to
    // BShamiltonian(basis)).
    std::vector<cdouble> HBS() const;
    long long hbsDim() const { return subsetDim_ > 0 ? subsetDim_ : excitonbasisdim; }
    std::vector<cdouble> HBS32() const;  // FP32 matrix (widened), when built in fp32
    bool hasHBS() const { return built_; }
    // Lattice Fourier stacks (ftMotifStack natoms x natoms x nmesh, ftMotifQ natoms x natoms).
    std::vector<cdouble> ftMotifStack() const;
    std::vector<cdouble> ftMotifQ() const;
    std::vector<int> cellCoefficients() const;  // nR x ndim
    // y = H x with the resident matrix (or matrix-free), x: N x nvec column-major.
    std::vector<cdouble> applyBSE(const std::vector<cdouble>& x, int nvec, const std::string& precision,
                                  bool matrixFree) const;
    // Hellmann-Feynman gradients of the given states with respect to "r0", "eps_m", "eps_s".
    std::vector<double> hfGradients(const std::vector<cdouble>& eigvec, int nstates,
                                    const std::vector<std::string>& params) const;
    // dE_n/dH(R) for every cell and orbital pair (nstates x ncells x basisdim x basisdim,
    // blocks column-major); eigvec empty: the last solve's vectors.
    std::vector<cdouble> hoppingGradients(const std::vector<cdouble>& eigvec, int nstates) const;
    xatu_timings timings() const;
    xatu_solve_report lastSolveReport() const { return lastReport_; }
    xatu_ctx* context() const { return ctx_; }

    // Basis helpers (Exciton.tpp:128-199)
    std::vector<long long> createBasis(const std::vector<long long>& conduction,
                                       const std::vector<long long>& valence) const;
    // Fix the phase of every column so that the sum of its coefficients is real >= 0
    // (ExcitonTB.cpp:673-683).
    static void fixGlobalPhase(std::vector<cdouble>& coefs, int rows, int cols);

private:
    friend class ResultTB;
    void initializeExcitonAttributes(int ncell, const std::vector<long long>& bands,
                                     const std::vector<double>& parameters, const std::vector<double>& Q);
    void initializeExcitonAttributes(const ExcitonConfiguration& cfg);
    void splitBands(const std::vector<long long>& bands);
    void setBandsFromCount(int nbands, int nrmbands);
    void initializeResultsH0();
    void ensureContext();
    void check(int rc) const;

    // band stacks in the gauge the BSE uses (a copy with the atomic-gauge phases, or the
    // lattice-gauge stacks themselves)
    const std::vector<cdouble>& bseStack(int which) const;
    std::vector<cdouble> atomicK_, atomicKQ_;

    xatu_ctx* ctx_ = nullptr;
    long long subsetDim_ = 0;  // > 0: the resident matrix is that of a subset basis
    bool built_ = false;
    bool built32_ = false;
    bool initialized_ = false;
    xatu_solve_report lastReport_{};
};

class ResultTB {
public:
    ResultTB(ExcitonTB* exciton, std::vector<double> eigval, std::vector<cdouble> eigvec, int ncols);
    ExcitonTB* exciton;
    std::vector<double> eigval;   // ascending
    std::vector<cdouble> eigvec;  // N x ncols column-major
    int ncols = 0;
    long long nrows = 0;

    // Observables (Result.tpp:16-60)
    double kineticEnergy(int stateindex) const;
    double potentialEnergy(int stateindex) const;
    // Optics (src/ResultTB.cpp:242-298, 412-460; src/skubo_w.f90): oscillator strengths
    // 3 x ncols (a.u.), and the Lorentzian-broadened exciton conductivity, nw x {w, xx, xy, yx, yy}.
    std::vector<cdouble> excitonOscillatorStrength() const;
    std::vector<double> absorptionSpectrum(double w0, double wrange, int nw, double eta) const;
    // Output formats (Result.tpp:404-448)
    void writeEigenvalues(FILE* textfile, int n = 0) const;
    void writeStates(FILE* textfile, int n = 0) const;
    // k-space amplitude (Result.tpp:163-190): per k, sum_pairs |coef|^2 / |k_1 - k_0|
    std::vector<double> reciprocalAmplitude(const std::vector<cdouble>& coefs) const;
    std::vector<double> reciprocalAmplitude(int stateindex) const;
    void writeReciprocalAmplitude(const std::vector<cdouble>& coefs, FILE* textfile) const;
    void writeReciprocalAmplitude(int stateindex, FILE* textfile) const;
    void writeExtendedReciprocalAmplitude(const std::vector<cdouble>& coefs, FILE* textfile) const;
    void writeExtendedReciprocalAmplitude(int stateindex, FILE* textfile) const;
    void writePhase(const std::vector<cdouble>& coefs, FILE* textfile) const;
    void writePhase(int stateindex, FILE* textfile) const;
    void writeExtendedPhase(const std::vector<cdouble>& coefs, FILE* textfile) const;
    void writeExtendedPhase(int stateindex, FILE* textfile) const;
    // Spin <S_z>: {total, hole, electron} (src/ResultTB.cpp:21-99, Result.tpp:145-150, 456-470)
    std::vector<cdouble> spinX(const std::vector<cdouble>& coefs) const;
    std::vector<cdouble> spinX(int stateindex) const;
    void writeSpin(int n, FILE* textfile) const;
    // Real-space electron probability with a fixed hole (src/ResultTB.cpp:310-411);
    // cells and positions are Cartesian 3-vectors.
    double realSpaceWavefunction(const std::vector<cdouble>& coefs, int electronIndex, int holeIndex,
                                 const std::vector<double>& eCell, const std::vector<double>& hCell) const;
    void writeRealspaceAmplitude(const std::vector<cdouble>& coefs, int holeIndex, const std::vector<double>& holeCell,
                                 FILE* textfile, int ncells) const;
    void writeRealspaceAmplitude(int stateindex, int holeIndex, const std::vector<double>& holeCell, FILE* textfile,
                                 int ncells) const;
    // Result.tpp:66-139, 479-499
    double bindingEnergy(int stateindex, double gap = -1) const;
    double determineGap() const;
    int findExcitonPeak(int stateindex) const;
    double boundingBoxBZ() const;
    double ftExcitonEnvelope(int stateindex, const std::vector<double>& electronPosition,
                             const std::vector<double>& holePosition) const;
};

// Single-particle velocity matrix elements of the k stack, V[((k*3 + j)*nb + n)*nb + m] (a.u.).
std::vector<cdouble> velocityMatrixElements(const ExcitonTB& ex);

// Degenerate-level grouping of sorted energies (src/utils.cpp:121-152).
std::vector<std::vector<double>> detectDegeneracies(const std::vector<double>& eigval, int n, int precision);

}  // namespace xatu
