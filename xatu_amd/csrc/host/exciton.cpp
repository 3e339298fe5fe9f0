#include "exciton.hpp"

#include <cmath>
#include <cstdio>
#include <iostream>
#include <stdexcept>

namespace xatu {

namespace {

int potential_id(const std::string& p) {
    if (p == "keldysh") return XATU_POTENTIAL_KELDYSH;
    if (p == "coulomb") return XATU_POTENTIAL_COULOMB;
    throw std::invalid_argument("selectPotential(): potential must be either 'keldysh' or 'coulomb'");
}

long long ipow(long long b, int e) {
    long long r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// construction (src/ExcitonTB.cpp:20-210)

void ExcitonTB::initializeExcitonAttributes(int ncell_, const std::vector<long long>& bands_,
                                            const std::vector<double>& parameters, const std::vector<double>& Q_) {
    if (parameters.size() != 3) throw std::invalid_argument("parameters array must be 3d (eps_m, eps_s, r0)");
    if (Q_.size() != 3) throw std::invalid_argument("Q vector must be 3d");
    ncell = ncell_;
    totalCells = ipow(ncell, system->ndim);
    Q = Q_;
    bands = bands_;
    eps_m = parameters[0];
    eps_s = parameters[1];
    r0 = parameters[2];
    cutoff = ncell / 2.5;  // default cutoff, ExcitonTB.cpp:29
    if (r0 == 0) throw std::invalid_argument("Error: r0 must be non-zero");
}

void ExcitonTB::splitBands(const std::vector<long long>& b) {
    valenceBands.clear();
    conductionBands.clear();
    for (long long x : b) {
        if (x <= 0) {
            valenceBands.push_back(x + system->fermiLevel);
        } else {
            conductionBands.push_back(x + system->fermiLevel);
        }
    }
    bandList = valenceBands;
    bandList.insert(bandList.end(), conductionBands.begin(), conductionBands.end());
}

void ExcitonTB::setBandsFromCount(int nb, int nrm) {
    const int f = system->fermiLevel;
    valenceBands.clear();
    conductionBands.clear();
    for (long long x = f - nb - nrm + 1; x <= f - nrm; ++x) valenceBands.push_back(x);
    for (long long x = f + 1 + nrm; x <= f + nb + nrm; ++x) conductionBands.push_back(x);
    bandList = valenceBands;
    bandList.insert(bandList.end(), conductionBands.begin(), conductionBands.end());
    bands.clear();
    for (long long x : bandList) bands.push_back(x - f);
    nbands = nb;
    nrmbands = nrm;
}

ExcitonTB::ExcitonTB(const SystemConfiguration& config, int ncell_, int nbands_, int nrmbands_,
                     const std::vector<double>& parameters, const std::vector<double>& Q_)
    : ExcitonTB(std::make_shared<SystemTB>(config), ncell_, nbands_, nrmbands_, parameters, Q_) {}

ExcitonTB::ExcitonTB(const SystemConfiguration& config, int ncell_, const std::vector<long long>& bands_,
                     const std::vector<double>& parameters, const std::vector<double>& Q_)
    : ExcitonTB(std::make_shared<SystemTB>(config), ncell_, bands_, parameters, Q_) {}

ExcitonTB::ExcitonTB(std::shared_ptr<SystemTB> sys, int ncell_, int nbands_, int nrmbands_,
                     const std::vector<double>& parameters, const std::vector<double>& Q_)
    : system(std::move(sys)) {
    initializeExcitonAttributes(ncell_, {}, parameters, Q_);
    if (2 * nbands_ > system->basisdim) {
        throw std::invalid_argument("Error: Number of bands cannot be higher than actual material bands");
    }
    setBandsFromCount(nbands_, nrmbands_);
}

ExcitonTB::ExcitonTB(std::shared_ptr<SystemTB> sys, int ncell_, const std::vector<long long>& bands_,
                     const std::vector<double>& parameters, const std::vector<double>& Q_)
    : system(std::move(sys)) {
    initializeExcitonAttributes(ncell_, bands_, parameters, Q_);
    if (static_cast<long long>(bands_.size()) > system->basisdim) {
        throw std::invalid_argument("Error: Number of bands cannot be higher than actual material bands");
    }
    splitBands(bands_);
}

ExcitonTB::ExcitonTB(const SystemConfiguration& config, const ExcitonConfiguration& cfg)
    : system(std::make_shared<SystemTB>(config)) {
    initializeExcitonAttributes(cfg);
}

// From an exciton file (ExcitonTB.cpp:42-80). As in the reference, the file's
// 'cutoff' key is parsed but not applied here.
void ExcitonTB::initializeExcitonAttributes(const ExcitonConfiguration& cfg) {
    const ExcitonInfo& e = cfg.excitonInfo;
    std::vector<long long> b = e.bands;
    if (b.empty()) {
        for (long long x = -e.nbands + 1; x <= e.nbands; ++x) b.push_back(x);
    }
    if (e.eps.size() < 3) throw std::invalid_argument("'dielectric' must give eps_m, eps_s and r0");
    initializeExcitonAttributes(e.ncell, b, {e.eps[0], e.eps[1], e.eps[2]}, e.Q);
    nbands = e.nbands;
    splitBands(b);
    exchange = e.exchange;
    scissor = e.scissor;
    mode = e.mode;
    nReciprocalVectors = e.nReciprocalVectors;
    regularization = e.regularization;
    if (regularization == 0) regularization = system->a;
    potential = e.potential;
    exchangePotential = e.exchangePotential;
}

ExcitonTB::~ExcitonTB() {
    if (ctx_) xatu_destroy(ctx_);
}

// ---------------------------------------------------------------------------
// setters (ExcitonTB.cpp:227-296, Exciton.tpp:15-117)

void ExcitonTB::setParameters(const std::vector<double>& p) {
    if (p.size() == 3) {
        eps_m = p[0];
        eps_s = p[1];
        r0 = p[2];
    } else {
        std::cout << "parameters array must be 3d (eps_m, eps_s, r0)" << std::endl;
    }
}
void ExcitonTB::setParameters(double em, double es, double r) {
    eps_m = em;
    eps_s = es;
    r0 = r;
}
void ExcitonTB::setGauge(const std::string& g) {
    if (g != "lattice" && g != "atomic") throw std::invalid_argument("setGauge(): gauge must be either lattice or atomic");
    gauge = g;
}
void ExcitonTB::setMode(const std::string& m) {
    if (m != "realspace" && m != "reciprocalspace") {
        throw std::invalid_argument("setMode(): mode must be either realspace or reciprocalspace");
    }
    mode = m;
}
void ExcitonTB::setReciprocalVectors(int n) {
    if (n < 0) throw std::invalid_argument("setReciprocalVectors(): given number must be positive");
    nReciprocalVectors = n;
}
void ExcitonTB::setRegularization(double r) { regularization = r; }
void ExcitonTB::setUnitCells(int n) {
    if (n > 0) {
        ncell = n;
    } else {
        std::cout << "ncell must be a positive number" << std::endl;
    }
}
void ExcitonTB::setBands(const std::vector<long long>& b) {
    bands = b;
    splitBands(b);
}
void ExcitonTB::setBands(int nb, int nrm) {
    if (nb > 0 && nrm > 0) {
        setBandsFromCount(nb, nrm);
    } else {
        std::cout << "Included bands and removed bands must be positive numbers" << std::endl;
    }
}
void ExcitonTB::setQ(const std::vector<double>& q) {
    if (q.size() == 3) {
        Q = q;
    } else {
        std::cout << "Q vector must be 3d" << std::endl;
    }
}
void ExcitonTB::setCutoff(double c) {
    if (c > 0) {
        cutoff = c;
        if (c > ncell) std::cout << "Warning: cutoff is higher than number of unit cells" << std::endl;
    } else {
        std::cout << "cutoff must be a positive number" << std::endl;
    }
}
void ExcitonTB::setScissor(double s) { scissor = s; }
void ExcitonTB::setExchange(bool x) { exchange = x; }
void ExcitonTB::setPotential(const std::string& p) {
    potential_id(p);
    potential = p;
}
void ExcitonTB::setExchangePotential(const std::string& p) {
    potential_id(p);
    exchangePotential = p;
}
void ExcitonTB::setPrecision(const std::string& p) {
    if (p != "fp64" && p != "fp32") throw std::invalid_argument("setPrecision(): precision must be fp64 or fp32");
    precision = p;
}
void ExcitonTB::setMatrixFree(bool m) { matrixFree = m; }
void ExcitonTB::setDavidsonBlock(int b) { davidsonBlock = b; }
void ExcitonTB::setDavidsonTolerance(double t) { davidsonTol = t; }
void ExcitonTB::setDevice(int d) {
    if (ctx_) throw std::logic_error("setDevice(): the device must be chosen before initializeHamiltonian()");
    device = d;
}

// ---------------------------------------------------------------------------

std::vector<long long> ExcitonTB::createBasis(const std::vector<long long>& conduction,
                                              const std::vector<long long>& valence) const {
    std::vector<long long> states;
    states.reserve(static_cast<size_t>(system->nk) * conduction.size() * valence.size() * 3);
    for (int i = 0; i < system->nk; ++i)
        for (size_t k = 0; k < conduction.size(); ++k)
            for (size_t j = 0; j < valence.size(); ++j) {
                states.push_back(valence[j]);
                states.push_back(conduction[k]);
                states.push_back(i);
            }
    return states;
}

void ExcitonTB::fixGlobalPhase(std::vector<cdouble>& coefs, int rows, int cols) {
    for (int j = 0; j < cols; ++j) {
        cdouble s(0.0, 0.0);
        for (int i = 0; i < rows; ++i) s += coefs[i + static_cast<size_t>(j) * rows];
        const double phase = std::arg(s);
        const cdouble f = std::exp(cdouble(0.0, -phase));
        for (int i = 0; i < rows; ++i) coefs[i + static_cast<size_t>(j) * rows] *= f;
    }
}

void ExcitonTB::brillouinZoneMesh(int n) { system->brillouinZoneMesh(n); }

void ExcitonTB::check(int rc) const {
    if (rc == XATU_OK) return;
    const std::string msg = xatu_last_error(ctx_);
    if (rc == XATU_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    if (rc == XATU_ERR_LOGIC) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

void ExcitonTB::ensureContext() {
    if (ctx_) return;
    const int rc = xatu_create(&ctx_, device);
    if (rc != XATU_OK) {
        ctx_ = nullptr;
        throw std::runtime_error("xatu_create failed (no usable HIP device?)");
    }
}

// ExcitonTB::initializeHamiltonian (src/ExcitonTB.cpp:794-815).
void ExcitonTB::initializeHamiltonian() {
    if (bands.empty()) throw std::invalid_argument("Error: Exciton object must have some bands");
    if (system->nk == 0) throw std::invalid_argument("Error: BZ mesh must be initialized first");
    if (valenceBands.empty() || conductionBands.empty()) {
        throw std::invalid_argument("Error: Exciton object must have valence and conduction bands");
    }
    for (long long b : bandList) {
        if (b < 0 || b >= system->basisdim) {
            throw std::invalid_argument("Error: Number of bands cannot be higher than actual material bands");
        }
    }
    potential_id(potential);
    if (exchange) potential_id(exchangePotential);
    if (regularization < 1E-10) regularization = system->a;
    excitonbasisdim = static_cast<long long>(system->nk) * valenceBands.size() * conductionBands.size();
    totalCells = ipow(static_cast<long long>(ncell) * system->factor, system->ndim);
    basisStates = createBasis(conductionBands, valenceBands);
    bandToIndex.clear();
    for (size_t i = 0; i < bandList.size(); ++i) bandToIndex[static_cast<int>(bandList[i])] = static_cast<int>(i);
    initializeResultsH0();
    initialized_ = true;
}

// Band stacks on the host, then the lattice sums on the device
// (ExcitonTB::initializeResultsH0, src/ExcitonTB.cpp:707-787).
void ExcitonTB::initializeResultsH0() {
    const int nk = system->nk;
    const int nb = static_cast<int>(bandList.size());
    const int bd = system->basisdim;
    eigvalKStack.assign(static_cast<size_t>(nb) * nk, 0.0);
    eigvalKQStack.assign(static_cast<size_t>(nb) * nk, 0.0);
    eigvecKStack.assign(static_cast<size_t>(bd) * nb * nk, cdouble(0.0, 0.0));
    eigvecKQStack.assign(static_cast<size_t>(bd) * nb * nk, cdouble(0.0, 0.0));
    const bool hasQ = (Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2]) != 0.0;

#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nk; ++i) {
        std::vector<double> w;
        std::vector<cdouble> v;
        auto solve_into = [&](const double* k, std::vector<double>& evals, std::vector<cdouble>& evecs) {
            system->solveBands(k, w, v);
            fixGlobalPhase(v, bd, bd);
            for (int b = 0; b < nb; ++b) {
                const long long col = bandList[b];
                evals[static_cast<size_t>(i) * nb + b] = w[col];
                for (int a = 0; a < bd; ++a)
                    evecs[(static_cast<size_t>(i) * nb + b) * bd + a] = v[a + static_cast<size_t>(col) * bd];
            }
        };
        const double* k = system->kpoints.data() + 3 * static_cast<size_t>(i);
        solve_into(k, eigvalKStack, eigvecKStack);
        if (hasQ) {
            const double kq[3] = {k[0] + Q[0], k[1] + Q[1], k[2] + Q[2]};
            solve_into(kq, eigvalKQStack, eigvecKQStack);
        }
    }
    if (!hasQ) {
        eigvalKQStack = eigvalKStack;
        eigvecKQStack = eigvecKStack;
    }

    ensureContext();
    std::vector<double> tau(static_cast<size_t>(system->natoms) * 3);
    for (int a = 0; a < system->natoms; ++a)
        for (int x = 0; x < 3; ++x) tau[3 * a + x] = system->motif[4 * a + x];
    // Atomic gauge (SystemTB::latticeToAtomicGauge, src/SystemTB.cpp:166-182, applied in
    // BShamiltonian, src/ExcitonTB.cpp:877-884): coefficient alpha of a state at k times
    // exp(-i k . tau_atom(alpha)). The k+Q stack takes the phase of k as well, as the
    // reference evaluates it at kpoints.row(k). The exposed stacks stay in the lattice gauge.
    atomicK_.clear();
    atomicKQ_.clear();
    if (gauge == "atomic") {
        const std::vector<int> opa = system->orbitalsPerAtom();
        atomicK_ = eigvecKStack;
        atomicKQ_ = eigvecKQStack;
        for (int i = 0; i < nk; ++i) {
            const double* k = system->kpoints.data() + 3 * static_cast<size_t>(i);
            int alpha = 0;
            for (int a = 0; a < system->natoms; ++a) {
                const double kt = k[0] * tau[3 * a] + k[1] * tau[3 * a + 1] + k[2] * tau[3 * a + 2];
                const cdouble ph(std::cos(kt), -std::sin(kt));
                for (int o = 0; o < opa[a]; ++o, ++alpha)
                    for (int b = 0; b < nb; ++b) {
                        const size_t idx = (static_cast<size_t>(i) * nb + b) * bd + alpha;
                        atomicK_[idx] *= ph;
                        atomicKQ_[idx] *= ph;
                    }
            }
        }
    }
    const std::vector<int> opa = system->orbitalsPerAtom();
    check(xatu_set_lattice(ctx_, system->ndim, system->bravaisLattice.data(), system->natoms, tau.data(), opa.data(),
                           system->meshPoints));
    check(xatu_set_bands(ctx_, nk, static_cast<int>(valenceBands.size()), static_cast<int>(conductionBands.size()),
                         eigvalKStack.data(), eigvalKQStack.data(),
                         reinterpret_cast<const double*>(bseStack(0).data()),
                         reinterpret_cast<const double*>(bseStack(1).data()), system->kpointOffsets.data()));
    xatu_interaction in{};
    in.eps_m = eps_m;
    in.eps_s = eps_s;
    in.r0 = r0;
    in.regularization = regularization;
    in.cutoff = cutoff;
    in.ncell = ncell;
    in.total_cells = static_cast<int>(totalCells);
    in.potential = potential_id(potential);
    in.exchange_potential = potential_id(exchange ? exchangePotential : potential);
    in.exchange = exchange ? 1 : 0;
    for (int x = 0; x < 3; ++x) in.Q[x] = Q[x];
    in.mode = (mode == "reciprocalspace") ? XATU_MODE_RECIPROCAL : XATU_MODE_REALSPACE;
    in.n_reciprocal = nReciprocalVectors;
    check(xatu_fourier_sums(ctx_, &in));
    built_ = built32_ = false;
    subsetDim_ = 0;
}

const std::vector<cdouble>& ExcitonTB::bseStack(int which) const {
    if (gauge == "atomic" && !atomicK_.empty()) return which ? atomicKQ_ : atomicK_;
    return which ? eigvecKQStack : eigvecKStack;
}

// ExcitonTB::BShamiltonian (src/ExcitonTB.cpp:822-921).
void ExcitonTB::BShamiltonian() {
    if (!initialized_) throw std::logic_error("BShamiltonian(): call initializeHamiltonian() first");
    std::cout << "BSE dimension: " << excitonbasisdim << std::endl;
    subsetDim_ = 0;
    if (precision == "fp32") {
        check(xatu_build_bse(ctx_, XATU_FP32, scissor, nullptr));
        built32_ = true;
    } else {
        check(xatu_build_bse(ctx_, XATU_FP64, scissor, nullptr));
        built_ = true;
    }
}

// ExcitonTB::BShamiltonian(const arma::imat& basis) (src/ExcitonTB.cpp:840-921): the BSE
// over the given rows {valence band, conduction band, k index} (absolute band indices, as
// in basisStates), built in FP64 and kept resident as the operator of diagonalize().
void ExcitonTB::BShamiltonian(const std::vector<long long>& basis) {
    if (basis.empty()) {
        BShamiltonian();
        return;
    }
    if (!initialized_) throw std::logic_error("BShamiltonian(): call initializeHamiltonian() first");
    if (basis.size() % 3 != 0) throw std::invalid_argument("BShamiltonian(basis): basis must have 3 columns");
    const long long n = static_cast<long long>(basis.size() / 3);
    std::vector<int> cols(basis.size());
    for (long long i = 0; i < n; ++i) {
        for (int j = 0; j < 2; ++j) {
            const auto it = bandToIndex.find(static_cast<int>(basis[3 * i + j]));
            if (it == bandToIndex.end()) {
                throw std::invalid_argument("BShamiltonian(basis): band " + std::to_string(basis[3 * i + j]) +
                                            " is not among the exciton's bands");
            }
            cols[3 * i + j] = it->second;
        }
        cols[3 * i + 2] = static_cast<int>(basis[3 * i + 2]);
    }
    std::cout << "BSE dimension: " << n << std::endl;
    check(xatu_build_bse_basis(ctx_, static_cast<int>(n), cols.data(), scissor, nullptr));
    built_ = true;
    built32_ = false;
    subsetDim_ = n;
}

std::unique_ptr<ResultTB> ExcitonTB::diagonalize(const std::string& method, int nstates) {
    if (!built_ && !built32_ && !(matrixFree && initialized_)) {
        throw std::invalid_argument("diagonalizeRaw(): BSE Hamiltonian is not initialized.");
    }
    xatu_solve_options opt{};
    const long long N = hbsDim();
    int ncols;
    if (method == "diag") {
        if (!built_) throw std::invalid_argument("diagonalize('diag') needs the FP64 matrix (setPrecision('fp64'))");
        opt.method = XATU_SOLVE_DIAG;
        opt.nstates = static_cast<int>(N);
        ncols = static_cast<int>(N);
    } else if (method == "davidson" || method == "sparse") {
        opt.method = XATU_SOLVE_DAVIDSON;
        opt.nstates = nstates;
        // a subset-basis Hamiltonian is solved with its resident FP64 matrix
        opt.precision = (precision == "fp32" && subsetDim_ == 0) ? XATU_FP32 : XATU_FP64;
        opt.matrix_free = (matrixFree && subsetDim_ == 0) ? 1 : 0;
        opt.block = davidsonBlock;
        opt.tol = davidsonTol;
        ncols = nstates;
    } else {
        throw std::invalid_argument("diagonalize(): method must be 'diag', 'davidson' or 'sparse'");
    }
    std::vector<double> eigval(ncols);
    std::vector<cdouble> eigvec(static_cast<size_t>(N) * ncols);
    xatu_solve_report rep{};
    check(xatu_solve(ctx_, &opt, eigval.data(), reinterpret_cast<double*>(eigvec.data()), &rep));
    lastReport_ = rep;
    return std::make_unique<ResultTB>(this, std::move(eigval), std::move(eigvec), ncols);
}

std::vector<cdouble> ExcitonTB::HBS() const {
    if (!built_) throw std::logic_error("HBS: the FP64 BSE matrix has not been built");
    std::vector<cdouble> h(static_cast<size_t>(hbsDim()) * hbsDim());
    check(xatu_get_bse(ctx_, XATU_FP64, h.data()));
    return h;
}

std::vector<cdouble> ExcitonTB::HBS32() const {
    if (!built32_) throw std::logic_error("HBS32: the FP32 BSE matrix has not been built");
    const size_t nn = static_cast<size_t>(excitonbasisdim) * excitonbasisdim;
    std::vector<std::complex<float>> h(nn);
    check(xatu_get_bse(ctx_, XATU_FP32, h.data()));
    std::vector<cdouble> out(nn);
    for (size_t i = 0; i < nn; ++i) out[i] = cdouble(h[i].real(), h[i].imag());
    return out;
}

std::vector<cdouble> ExcitonTB::ftMotifStack() const {
    if (!initialized_) throw std::logic_error("ftMotifStack: call initializeHamiltonian() first");
    const size_t na = system->natoms;
    size_t nmesh = 1;
    for (int j = 0; j < system->ndim; ++j) nmesh *= system->meshPoints;
    std::vector<cdouble> out(na * na * nmesh);
    check(xatu_get_fourier(ctx_, reinterpret_cast<double*>(out.data()), nullptr));
    return out;
}

std::vector<cdouble> ExcitonTB::ftMotifQ() const {
    if (!initialized_) throw std::logic_error("ftMotifQ: call initializeHamiltonian() first");
    const size_t na = system->natoms;
    std::vector<cdouble> out(na * na);
    check(xatu_get_fourier(ctx_, nullptr, reinterpret_cast<double*>(out.data())));
    return out;
}

std::vector<int> ExcitonTB::cellCoefficients() const {
    int n = 0;
    check(xatu_get_cells(ctx_, &n, nullptr));
    std::vector<int> out(static_cast<size_t>(n) * system->ndim);
    check(xatu_get_cells(ctx_, &n, out.data()));
    return out;
}

std::vector<cdouble> ExcitonTB::applyBSE(const std::vector<cdouble>& x, int nvec, const std::string& prec,
                                         bool mf) const {
    if (static_cast<long long>(x.size()) != excitonbasisdim * nvec) throw std::invalid_argument("applyBSE: bad x");
    std::vector<cdouble> y(x.size());
    check(xatu_bse_matvec(ctx_, prec == "fp32" ? XATU_FP32 : XATU_FP64, mf ? 1 : 0, nvec,
                          reinterpret_cast<const double*>(x.data()), reinterpret_cast<double*>(y.data())));
    return y;
}

std::vector<double> ExcitonTB::hfGradients(const std::vector<cdouble>& eigvec, int nstates,
                                           const std::vector<std::string>& params) const {
    std::vector<int> ids;
    for (const auto& p : params) {
        if (p == "r0") {
            ids.push_back(XATU_PARAM_R0);
        } else if (p == "eps_m") {
            ids.push_back(XATU_PARAM_EPS_M);
        } else if (p == "eps_s") {
            ids.push_back(XATU_PARAM_EPS_S);
        } else {
            throw std::invalid_argument("hfGradients(): unknown parameter " + p);
        }
    }
    std::vector<double> g(ids.size() * static_cast<size_t>(nstates));
    const double* v = eigvec.empty() ? nullptr : reinterpret_cast<const double*>(eigvec.data());
    check(xatu_hf_gradients(ctx_, nstates, v, static_cast<int>(ids.size()), ids.data(), g.data()));
    return g;
}

// Gradients of exciton energies with respect to the tight-binding matrices H(R)
// (no reference counterpart). The device returns, per state, the sensitivities to the
// band data (w = dE/d eps_b(k), g = dE/d conj C_b(k)); first-order perturbation theory
// of the Bloch Hamiltonian, d eps_n = C_n^H dh C_n and dC_n = sum_{m != n} C_m
// (C_m^H dh C_n) / (eps_n - eps_m), turns them into dE = Re Tr(A(k) dh(k)) with
//   A(k) = sum_n [ w_n C_n C_n^H + 2 sum_{m != n} (g_n^H C_m) / (eps_n - eps_m) C_n C_m^H ],
// n over the bands the BSE uses at that k (valence of the k stack, conduction of the
// k+Q stack), m over all bands of h(k). The phase of C_n cancels from E, so the gauge
// term of dC_n is left out. With h(k) = sum_R H(R) e^{i k.R}, G(R)_ab = sum_k A(k)_ba
// e^{i k.R} gives dE = Re sum_R sum_ab G(R)_ab dH(R)_ab for any change of the H(R)
// that keeps h(k) Hermitian. Output: nstates x ncells x basisdim x basisdim complex,
// each block column-major like hamiltonianMatrices.
std::vector<cdouble> ExcitonTB::hoppingGradients(const std::vector<cdouble>& eigvec, int nstates) const {
    if (!system->overlapMatrices.empty() || system->isTriangular) {
        throw std::invalid_argument(
            "hoppingGradients(): needs an orthonormal basis with complete (not triangular) H(R) matrices");
    }
    const int nk = system->nk;
    const int nb = static_cast<int>(bandList.size());
    const int nv = static_cast<int>(valenceBands.size());
    const int bd = system->basisdim;
    const int nR = system->ncells;
    std::vector<double> w(static_cast<size_t>(nstates) * 2 * nb * nk);
    std::vector<cdouble> g(static_cast<size_t>(nstates) * 2 * bd * nb * nk);
    const double* v = eigvec.empty() ? nullptr : reinterpret_cast<const double*>(eigvec.data());
    check(xatu_band_gradients(ctx_, nstates, v, w.data(), reinterpret_cast<double*>(g.data())));
    if (gauge == "atomic") {
        // the device differentiates with respect to the atomic-gauge coefficients
        // C_at = p C, p = exp(-i k.tau): dE/d conj(C) = conj(p) dE/d conj(C_at)
        const std::vector<int> opa = system->orbitalsPerAtom();
        for (int k = 0; k < nk; ++k) {
            const double* kp = system->kpoints.data() + 3 * static_cast<size_t>(k);
            int alpha = 0;
            for (int a = 0; a < system->natoms; ++a) {
                const double* t = system->motif.data() + 4 * static_cast<size_t>(a);
                const double kt = kp[0] * t[0] + kp[1] * t[1] + kp[2] * t[2];
                const cdouble cp(std::cos(kt), std::sin(kt));
                for (int o = 0; o < opa[a]; ++o, ++alpha)
                    for (int sst = 0; sst < 2 * nstates; ++sst)
                        for (int b = 0; b < nb; ++b) g[((static_cast<size_t>(sst) * nk + k) * nb + b) * bd + alpha] *= cp;
            }
        }
    }

    const size_t bd2 = static_cast<size_t>(bd) * bd;
    // A[st][s][k] (bd x bd, column-major)
    std::vector<cdouble> A(static_cast<size_t>(2) * nstates * nk * bd2, cdouble(0.0, 0.0));
    std::string err;
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nk; ++k) {
        std::vector<double> ev;
        std::vector<cdouble> W;
        for (int st = 0; st < 2; ++st) {
            double kv[3];
            for (int x = 0; x < 3; ++x) kv[x] = system->kpoints[3 * static_cast<size_t>(k) + x] + (st ? Q[x] : 0.0);
            system->solveBands(kv, ev, W);
            const std::vector<cdouble>& stackV = st ? eigvecKQStack : eigvecKStack;
            const std::vector<double>& stackE = st ? eigvalKQStack : eigvalKStack;
            const int b0 = st ? nv : 0, b1 = st ? nb : nv;
            for (int s = 0; s < nstates; ++s) {
                cdouble* Ak = A.data() + ((static_cast<size_t>(st) * nstates + s) * nk + k) * bd2;
                for (int b = b0; b < b1; ++b) {
                    const cdouble* Cn = stackV.data() + (static_cast<size_t>(k) * nb + b) * bd;
                    const double en = stackE[static_cast<size_t>(k) * nb + b];
                    const long long nabs = bandList[b];
                    const double wn = w[((static_cast<size_t>(s) * 2 + st) * nk + k) * nb + b];
                    const cdouble* gn = g.data() + (((static_cast<size_t>(s) * 2 + st) * nk + k) * nb + b) * bd;
                    // B = w_n C_n + 2 sum_{m != n} coef_m C_m, so that A += C_n B^H ... written as
                    // A_ab += C_n[a] conj(Bv[b]) with Bv = w_n C_n + 2 sum_m conj(coef_m) C_m
                    std::vector<cdouble> Bv(bd);
                    for (int a = 0; a < bd; ++a) Bv[a] = wn * Cn[a];
                    for (int m = 0; m < bd; ++m) {
                        if (m == nabs) continue;
                        const double de = en - ev[m];
                        if (std::abs(de) < 1e-9) {
#pragma omp critical
                            err = "hoppingGradients(): degenerate bands at a k-point; the gradient is not defined";
                            continue;
                        }
                        const cdouble* Cm = W.data() + static_cast<size_t>(m) * bd;
                        cdouble gc(0.0, 0.0);
                        for (int a = 0; a < bd; ++a) gc += std::conj(gn[a]) * Cm[a];
                        const cdouble coef = 2.0 * gc / de;
                        for (int a = 0; a < bd; ++a) Bv[a] += std::conj(coef) * Cm[a];
                    }
                    for (int bcol = 0; bcol < bd; ++bcol) {
                        const cdouble cb = std::conj(Bv[bcol]);
                        for (int a = 0; a < bd; ++a) Ak[a + static_cast<size_t>(bcol) * bd] += Cn[a] * cb;
                    }
                }
            }
        }
    }
    if (!err.empty()) throw std::invalid_argument(err);

    // G_s(R)_ab = sum_st sum_k A_st,s(k)_ba exp(i (k + st Q).R)
    std::vector<cdouble> G(static_cast<size_t>(nstates) * nR * bd2, cdouble(0.0, 0.0));
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < nstates; ++s) {
        for (int r = 0; r < nR; ++r) {
            cdouble* Gr = G.data() + (static_cast<size_t>(s) * nR + r) * bd2;
            const double* R = system->unitCellList.data() + 3 * static_cast<size_t>(r);
            for (int st = 0; st < 2; ++st) {
                for (int k = 0; k < nk; ++k) {
                    double ph = 0.0;
                    for (int x = 0; x < 3; ++x) ph += (system->kpoints[3 * static_cast<size_t>(k) + x] + (st ? Q[x] : 0.0)) * R[x];
                    const cdouble e(std::cos(ph), std::sin(ph));
                    const cdouble* Ak = A.data() + ((static_cast<size_t>(st) * nstates + s) * nk + k) * bd2;
                    for (int b = 0; b < bd; ++b)
                        for (int a = 0; a < bd; ++a) Gr[a + static_cast<size_t>(b) * bd] += Ak[b + static_cast<size_t>(a) * bd] * e;
                }
            }
        }
    }
    return G;
}

xatu_timings ExcitonTB::timings() const {
    xatu_timings t{};
    if (ctx_) xatu_get_timings(ctx_, &t);
    return t;
}

// ---------------------------------------------------------------------------
// ResultTB (src/ResultTB.cpp:11-12, include/xatu/Result.tpp:10-60, 404-448)

ResultTB::ResultTB(ExcitonTB* ex, std::vector<double> ev, std::vector<cdouble> vec, int nc)
    : exciton(ex), eigval(std::move(ev)), eigvec(std::move(vec)), ncols(nc) {
    nrows = exciton->excitonbasisdim;
}

double ResultTB::kineticEnergy(int s) const {
    if (s < 0 || s >= ncols) throw std::invalid_argument("kineticEnergy: state index out of range");
    const int nb = static_cast<int>(exciton->bandList.size());
    double e = 0.0;
    for (long long n = 0; n < nrows; ++n) {
        const int v = exciton->bandToIndex.at(static_cast<int>(exciton->basisStates[3 * n + 0]));
        const int c = exciton->bandToIndex.at(static_cast<int>(exciton->basisStates[3 * n + 1]));
        const long long k = exciton->basisStates[3 * n + 2];
        const double hk = exciton->eigvalKQStack[k * nb + c] - exciton->eigvalKStack[k * nb + v];
        e += std::norm(eigvec[n + static_cast<size_t>(s) * nrows]) * hk;
    }
    return e;
}

double ResultTB::potentialEnergy(int s) const { return eigval.at(s) - kineticEnergy(s); }

void ResultTB::writeEigenvalues(FILE* f, int n) const {
    if (n > nrows || n < 0) {
        throw std::invalid_argument("Optional argument n must be a positive integer equal or below basisdim");
    }
    fprintf(f, "%lld\t", nrows);
    const int m = (n == 0) ? static_cast<int>(eigval.size()) : std::min<int>(n, static_cast<int>(eigval.size()));
    for (int i = 0; i < m; ++i) fprintf(f, "%11.7lf\t", eigval[i]);
    fprintf(f, "\n");
}

void ResultTB::writeStates(FILE* f, int n) const {
    if (n > nrows || n < 0) {
        throw std::invalid_argument("Optional argument n must be a positive integer equal or below basisdim");
    }
    fprintf(f, "%lld\n", nrows);
    for (long long i = 0; i < nrows; ++i) {
        const long long k = exciton->basisStates[3 * i + 2];
        const double* kp = exciton->system->kpoints.data() + 3 * k;
        fprintf(f, "%11.7lf\t%11.7lf\t%11.7lf\t%d\t%d\n", kp[0], kp[1], kp[2],
                static_cast<int>(exciton->basisStates[3 * i + 0]), static_cast<int>(exciton->basisStates[3 * i + 1]));
    }
    const int m = (n == 0) ? ncols : std::min(n, ncols);
    for (int s = 0; s < m; ++s) {
        for (long long j = 0; j < nrows; ++j) {
            const cdouble v = eigvec[j + static_cast<size_t>(s) * nrows];
            fprintf(f, "%11.7lf\t%11.7lf\t", v.real(), v.imag());
        }
        fprintf(f, "\n");
    }
}

// Squared k-space amplitude of a state: sum over its band pairs at each k of |coef|^2,
// divided by |k_1 - k_0| (include/xatu/Result.tpp:163-177).
std::vector<double> ResultTB::reciprocalAmplitude(const std::vector<cdouble>& coefs) const {
    const SystemTB& s = *exciton->system;
    const long long nbc = static_cast<long long>(exciton->conductionBands.size() * exciton->valenceBands.size());
    if (static_cast<long long>(coefs.size()) != nrows) throw std::invalid_argument("reciprocalAmplitude: wrong state size");
    if (s.nk < 2) throw std::invalid_argument("reciprocalAmplitude: needs at least two k-points");
    double dk = 0.0;
    for (int x = 0; x < 3; ++x) dk += (s.kpoints[3 + x] - s.kpoints[x]) * (s.kpoints[3 + x] - s.kpoints[x]);
    dk = std::sqrt(dk);
    std::vector<double> amp(s.nk, 0.0);
    for (int k = 0; k < s.nk; ++k) {
        double a = 0.0;
        for (long long b = 0; b < nbc; ++b) a += std::norm(coefs[nbc * k + b]);
        amp[k] = a / dk;
    }
    return amp;
}

std::vector<double> ResultTB::reciprocalAmplitude(int stateindex) const {
    if (stateindex < 0 || stateindex >= ncols) throw std::invalid_argument("reciprocalAmplitude: state index out of range");
    std::vector<cdouble> c(eigvec.begin() + static_cast<size_t>(stateindex) * nrows,
                           eigvec.begin() + static_cast<size_t>(stateindex + 1) * nrows);
    return reciprocalAmplitude(c);
}

void ResultTB::writeReciprocalAmplitude(const std::vector<cdouble>& coefs, FILE* f) const {
    const SystemTB& s = *exciton->system;
    const std::vector<double> amp = reciprocalAmplitude(coefs);
    fprintf(f, "kx\tky\tkz\tProb.\n");
    for (int k = 0; k < s.nk; ++k) {
        const double* kp = s.kpoints.data() + 3 * static_cast<size_t>(k);
        fprintf(f, "%11.8lf\t%11.8lf\t%11.8lf\t%11.8lf\n", kp[0], kp[1], kp[2], amp[k]);
    }
    fprintf(f, "#\n");
}

void ResultTB::writeReciprocalAmplitude(int stateindex, FILE* f) const {
    if (stateindex < 0 || stateindex >= ncols) throw std::invalid_argument("writeReciprocalAmplitude: state index out of range");
    std::vector<cdouble> c(eigvec.begin() + static_cast<size_t>(stateindex) * nrows,
                           eigvec.begin() + static_cast<size_t>(stateindex + 1) * nrows);
    writeReciprocalAmplitude(c, f);
}

std::vector<std::vector<double>> detectDegeneracies(const std::vector<double>& eigval, int n, int precision) {
    if (n < 0) throw std::invalid_argument("detectDegeneracies: n must be a positive integer");
    if (n > static_cast<int>(eigval.size())) {
        throw std::invalid_argument("detectDegeneracies: n must be lower than total number of eigenstates");
    }
    std::vector<std::vector<double>> pairs;
    double prev = eigval[0];
    int deg = 1;
    const double threshold = std::pow(10, -precision);
    for (int i = 1; i < n; ++i) {
        const double e = eigval[i];
        if (std::abs(e - prev) < threshold) {
            deg++;
        } else {
            pairs.push_back({prev, static_cast<double>(deg)});
            deg = 1;
        }
        prev = e;
    }
    pairs.push_back({prev, static_cast<double>(deg)});
    return pairs;
}

}  // namespace xatu
