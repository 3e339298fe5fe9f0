// Python bindings of the C++ facade (used by the test-suite and bench). Names and
// argument meanings follow the reference C++ API; matrices come back as numpy arrays
// with the reference's (Armadillo, column-major) shapes.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdio>
#include <cstring>

#include "../common/heev.hpp"
#include "crystal.hpp"
#include "exciton.hpp"

namespace py = pybind11;
using namespace xatu;

namespace {

py::array_t<double> rows(const std::vector<double>& v, int ncol) {
    const py::ssize_t n = ncol ? static_cast<py::ssize_t>(v.size() / ncol) : 0;
    py::array_t<double> a({n, static_cast<py::ssize_t>(ncol)});
    if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(double));
    return a;
}

py::array_t<long long> irows(const std::vector<long long>& v, int ncol) {
    const py::ssize_t n = static_cast<py::ssize_t>(v.size() / ncol);
    py::array_t<long long> a({n, static_cast<py::ssize_t>(ncol)});
    if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(long long));
    return a;
}

template <typename T>
py::array_t<T> fortran(const std::vector<T>& v, std::vector<py::ssize_t> shape) {
    py::array_t<T, py::array::f_style> a(shape);
    if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
    return a;
}

std::vector<cdouble> to_vec(py::array_t<cdouble, py::array::f_style | py::array::forcecast> a) {
    return std::vector<cdouble>(a.data(), a.data() + a.size());
}

FILE* open_or_throw(const std::string& name) {
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) throw std::runtime_error("cannot open " + name);
    return f;
}

}  // namespace

PYBIND11_MODULE(_xatu, m) {
    m.doc() = "MI355X BSE kernel builder and exciton solver: ExcitonTB facade over the C-ABI";

    py::class_<SystemConfiguration>(m, "SystemConfiguration")
        .def(py::init<const std::string&>(), py::arg("filename"))
        .def_readonly("foundArguments", &SystemConfiguration::foundArguments)
        .def_property_readonly("ndim", [](const SystemConfiguration& c) { return c.systemInfo.ndim; })
        .def_property_readonly("filling", [](const SystemConfiguration& c) { return c.systemInfo.filling; })
        .def_property_readonly("norbitals", [](const SystemConfiguration& c) { return c.systemInfo.norbitals; })
        .def_property_readonly("bravaisLattice",
                               [](const SystemConfiguration& c) { return rows(c.systemInfo.bravaisLattice, 3); })
        .def_property_readonly("motif", [](const SystemConfiguration& c) { return rows(c.systemInfo.motif, 4); })
        .def_property_readonly("bravaisVectors",
                               [](const SystemConfiguration& c) { return rows(c.systemInfo.bravaisVectors, 3); })
        .def_property_readonly("hamiltonian",
                               [](const SystemConfiguration& c) {
                                   const py::ssize_t d = c.systemInfo.matrixDim;
                                   return fortran(c.systemInfo.hamiltonian, {d, d, c.systemInfo.nHamiltonian});
                               })
        .def_property_readonly("overlap", [](const SystemConfiguration& c) {
            const py::ssize_t d = c.systemInfo.matrixDim;
            return fortran(c.systemInfo.overlap, {d, d, c.systemInfo.nOverlap});
        });

    py::class_<ExcitonConfiguration>(m, "ExcitonConfiguration")
        .def(py::init<const std::string&>(), py::arg("filename"))
        .def_property_readonly("label", [](const ExcitonConfiguration& c) { return c.excitonInfo.label; })
        .def_property_readonly("ncell", [](const ExcitonConfiguration& c) { return c.excitonInfo.ncell; })
        .def_property_readonly("nbands", [](const ExcitonConfiguration& c) { return c.excitonInfo.nbands; })
        .def_property_readonly("submeshFactor", [](const ExcitonConfiguration& c) { return c.excitonInfo.submeshFactor; })
        .def_property_readonly("bands", [](const ExcitonConfiguration& c) { return c.excitonInfo.bands; })
        .def_property_readonly("Q", [](const ExcitonConfiguration& c) { return c.excitonInfo.Q; })
        .def_property_readonly("shift", [](const ExcitonConfiguration& c) { return c.excitonInfo.shift; })
        .def_property_readonly("cutoff", [](const ExcitonConfiguration& c) { return c.excitonInfo.cutoff; })
        .def_property_readonly("eps", [](const ExcitonConfiguration& c) { return c.excitonInfo.eps; })
        .def_property_readonly("mode", [](const ExcitonConfiguration& c) { return c.excitonInfo.mode; })
        .def_property_readonly("exchange", [](const ExcitonConfiguration& c) { return c.excitonInfo.exchange; })
        .def_property_readonly("scissor", [](const ExcitonConfiguration& c) { return c.excitonInfo.scissor; })
        .def_property_readonly("nReciprocalVectors",
                               [](const ExcitonConfiguration& c) { return c.excitonInfo.nReciprocalVectors; })
        .def_property_readonly("potential", [](const ExcitonConfiguration& c) { return c.excitonInfo.potential; })
        .def_property_readonly("exchangePotential",
                               [](const ExcitonConfiguration& c) { return c.excitonInfo.exchangePotential; })
        .def_property_readonly("regularization",
                               [](const ExcitonConfiguration& c) { return c.excitonInfo.regularization; });

    py::class_<CRYSTALConfiguration, SystemConfiguration>(m, "CRYSTALConfiguration")
        .def(py::init<const std::string&, int>(), py::arg("file"), py::arg("ncells") = 20)
        .def_readonly("natoms", &CRYSTALConfiguration::natoms)
        .def_readonly("norbitalsAO", &CRYSTALConfiguration::norbitals)
        .def_readonly("nsh", &CRYSTALConfiguration::nsh)
        .def_readonly("nspecies", &CRYSTALConfiguration::nspecies)
        .def_readonly("totalElectrons", &CRYSTALConfiguration::totalElectrons)
        .def_readonly("coreElectrons", &CRYSTALConfiguration::coreElectrons)
        .def_readonly("SOC_FLAG", &CRYSTALConfiguration::SOC_FLAG)
        .def_readonly("MAGNETIC_FLAG", &CRYSTALConfiguration::MAGNETIC_FLAG)
        .def_readonly("shellsPerSpecies", &CRYSTALConfiguration::shellsPerSpecies)
        .def_readonly("orbitalsPerSpecies", &CRYSTALConfiguration::orbitalsPerSpecies)
        .def_readonly("atomic_number_ordering", &CRYSTALConfiguration::atomic_number_ordering)
        .def_readonly("gaussianCoefficients", &CRYSTALConfiguration::gaussianCoefficients)
        .def_readonly("nOverlap", &CRYSTALConfiguration::nOverlap)
        .def_readonly("nFock", &CRYSTALConfiguration::nFock);

    py::class_<SystemTB, std::shared_ptr<SystemTB>>(m, "SystemTB")
        .def(py::init<const SystemConfiguration&>())
        .def_readonly("ndim", &SystemTB::ndim)
        .def_readonly("natoms", &SystemTB::natoms)
        .def_readonly("ncells", &SystemTB::ncells)
        .def_readonly("nk", &SystemTB::nk)
        .def_readonly("factor", &SystemTB::factor)
        .def_readonly("a", &SystemTB::a)
        .def_readonly("c", &SystemTB::c)
        .def_readonly("unitCellArea", &SystemTB::unitCellArea)
        .def_readonly("basisdim", &SystemTB::basisdim)
        .def_readonly("filling", &SystemTB::filling)
        .def_readonly("fermiLevel", &SystemTB::fermiLevel)
        .def_readonly("isTriangular", &SystemTB::isTriangular)
        .def_readonly("meshPoints", &SystemTB::meshPoints)
        .def_property_readonly("orbitals", [](const SystemTB& s) { return s.orbitals; })
        .def_property_readonly("bravaisLattice", [](const SystemTB& s) { return rows(s.bravaisLattice, 3); })
        .def_property_readonly("reciprocalLattice", [](const SystemTB& s) { return rows(s.reciprocalLattice, 3); })
        .def_property_readonly("motif", [](const SystemTB& s) { return rows(s.motif, 4); })
        .def_property_readonly("unitCellList", [](const SystemTB& s) { return rows(s.unitCellList, 3); })
        .def_property_readonly("kpoints", [](const SystemTB& s) { return rows(s.kpoints, 3); })
        .def_property_readonly("meshBZ", [](const SystemTB& s) { return rows(s.meshBZ, 3); })
        .def_property_readonly("kpointOffsets",
                               [](const SystemTB& s) {
                                   std::vector<long long> v(s.kpointOffsets.begin(), s.kpointOffsets.end());
                                   return irows(v, s.ndim);
                               })
        .def("orbitalsPerAtom", &SystemTB::orbitalsPerAtom)
        .def("brillouinZoneMesh", &SystemTB::brillouinZoneMesh)
        .def("reducedBrillouinZoneMesh", &SystemTB::reducedBrillouinZoneMesh)
        .def("shiftBZ", &SystemTB::shiftBZ)
        .def("setFilling", &SystemTB::setFilling)
        .def("setAU", &SystemTB::setAU)
        .def("setTriangular", &SystemTB::setTriangular)
        .def("calculateInverseReciprocalMatrix", &SystemTB::calculateInverseReciprocalMatrix)
        .def("findEquivalentPointBZ",
             [](const SystemTB& s, std::vector<double> k, int ncell) { return s.findEquivalentPointBZ(k.data(), ncell); })
        .def("truncateSupercell", [](const SystemTB& s, int ncell, double radius) {
            return rows(s.truncateSupercell(ncell, radius), 3);
        })
        .def("hamiltonian",
             [](const SystemTB& s, std::vector<double> k) {
                 return fortran(s.hamiltonian(k.data()), {s.basisdim, s.basisdim});
             })
        .def("solveBands", [](const SystemTB& s, std::vector<double> k) {
            std::vector<double> w;
            std::vector<cdouble> v;
            s.solveBands(k.data(), w, v);
            return py::make_tuple(py::array_t<double>(w.size(), w.data()), fortran(v, {s.basisdim, s.basisdim}));
        });

    py::class_<xatu_timings>(m, "Timings")
        .def_readonly("ms_fourier", &xatu_timings::ms_fourier)
        .def_readonly("ms_build", &xatu_timings::ms_build)
        .def_readonly("ms_build_kernel", &xatu_timings::ms_build_kernel)
        .def_readonly("ms_matvec_kernel", &xatu_timings::ms_matvec_kernel)
        .def_readonly("n_matvec_kernel", &xatu_timings::n_matvec_kernel)
        .def_readonly("n_matvec_vectors", &xatu_timings::n_matvec_vectors);

    py::class_<xatu_solve_report>(m, "SolveReport")
        .def_readonly("iterations", &xatu_solve_report::iterations)
        .def_readonly("products", &xatu_solve_report::products)
        .def_readonly("restarts", &xatu_solve_report::restarts)
        .def_readonly("refine_iterations", &xatu_solve_report::refine_iterations)
        .def_readonly("converged", &xatu_solve_report::converged)
        .def_readonly("max_residual", &xatu_solve_report::max_residual)
        .def_readonly("ms_total", &xatu_solve_report::ms_total)
        .def_readonly("ms_products", &xatu_solve_report::ms_products)
        .def_readonly("ms_rayleigh_ritz", &xatu_solve_report::ms_rayleigh_ritz)
        .def_readonly("ms_projections", &xatu_solve_report::ms_projections);

    auto result = py::class_<ResultTB>(m, "ResultTB");
    result
        .def_property_readonly("eigval", [](const ResultTB& r) { return py::array_t<double>(r.eigval.size(), r.eigval.data()); })
        .def_property_readonly("eigvec", [](const ResultTB& r) { return fortran(r.eigvec, {r.nrows, r.ncols}); })
        .def("excitonOscillatorStrength",
             [](const ResultTB& r) { return fortran(r.excitonOscillatorStrength(), {3, r.ncols}); })
        .def("absorptionSpectrum",
             [](const ResultTB& r, double w0, double wrange, int nw, double eta) {
                 std::vector<double> a = r.absorptionSpectrum(w0, wrange, nw, eta);
                 py::array_t<double> out({static_cast<py::ssize_t>(nw), static_cast<py::ssize_t>(5)});
                 std::copy(a.begin(), a.end(), out.mutable_data());
                 return out;
             },
             py::arg("w0") = 0.0, py::arg("wrange") = 6.0, py::arg("nw") = 2000, py::arg("eta") = 0.05)
        .def("kineticEnergy", &ResultTB::kineticEnergy)
        .def("potentialEnergy", &ResultTB::potentialEnergy)
        .def("writeEigenvalues",
             [](const ResultTB& r, const std::string& name, int n) {
                 FILE* f = open_or_throw(name);
                 r.writeEigenvalues(f, n);
                 std::fclose(f);
             },
             py::arg("filename"), py::arg("n") = 0)
        .def("reciprocalAmplitude",
             [](const ResultTB& r, int i) {
                 std::vector<double> a = r.reciprocalAmplitude(i);
                 return py::array_t<double>(a.size(), a.data());
             })
        .def("writeReciprocalAmplitude",
             [](const ResultTB& r, int i, const std::string& name) {
                 FILE* f = open_or_throw(name);
                 r.writeReciprocalAmplitude(i, f);
                 std::fclose(f);
             },
             py::arg("stateindex"), py::arg("filename"))
        .def("writeStates",
             [](const ResultTB& r, const std::string& name, int n) {
                 FILE* f = open_or_throw(name);
                 r.writeStates(f, n);
                 std::fclose(f);
             },
             py::arg("filename"), py::arg("n") = 0)
        .def("spinX",
             [](const ResultTB& r, py::object state) {
                 std::vector<cdouble> sp = py::isinstance<py::int_>(state)
                                               ? r.spinX(state.cast<int>())
                                               : r.spinX(state.cast<std::vector<cdouble>>());
                 return py::array_t<cdouble>(sp.size(), sp.data());
             },
             py::arg("state"))
        .def("writeSpin",
             [](const ResultTB& r, int n, const std::string& name) {
                 FILE* f = open_or_throw(name);
                 try {
                     r.writeSpin(n, f);
                 } catch (...) {
                     std::fclose(f);
                     throw;
                 }
                 std::fclose(f);
             },
             py::arg("n"), py::arg("filename"))
        .def("realSpaceWavefunction",
             [](const ResultTB& r, const std::vector<cdouble>& coefs, int e, int h, const std::vector<double>& ec,
                const std::vector<double>& hc) { return r.realSpaceWavefunction(coefs, e, h, ec, hc); },
             py::arg("coefs"), py::arg("electronIndex"), py::arg("holeIndex"), py::arg("eCell"), py::arg("hCell"))
        .def("writeRealspaceAmplitude",
             [](const ResultTB& r, py::object state, int hole, const std::vector<double>& hcell,
                const std::string& name, int ncells) {
                 FILE* f = open_or_throw(name);
                 try {
                     if (py::isinstance<py::int_>(state)) r.writeRealspaceAmplitude(state.cast<int>(), hole, hcell, f, ncells);
                     else r.writeRealspaceAmplitude(state.cast<std::vector<cdouble>>(), hole, hcell, f, ncells);
                 } catch (...) {
                     std::fclose(f);
                     throw;
                 }
                 std::fclose(f);
             },
             py::arg("state"), py::arg("holeIndex"), py::arg("holeCell"), py::arg("filename"), py::arg("ncells"))
        .def("bindingEnergy", &ResultTB::bindingEnergy, py::arg("stateindex"), py::arg("gap") = -1.0)
        .def("determineGap", &ResultTB::determineGap)
        .def("findExcitonPeak", &ResultTB::findExcitonPeak)
        .def("boundingBoxBZ", &ResultTB::boundingBoxBZ)
        .def("ftExcitonEnvelope", &ResultTB::ftExcitonEnvelope, py::arg("stateindex"), py::arg("electron_position"),
             py::arg("hole_position"));
    auto def_state_writer = [&](const char* name, void (ResultTB::*by_coefs)(const std::vector<cdouble>&, FILE*) const,
                                void (ResultTB::*by_index)(int, FILE*) const) {
        result.def(name, [by_coefs, by_index](const ResultTB& r, py::object state, const std::string& fname) {
            FILE* f = open_or_throw(fname);
            try {
                if (py::isinstance<py::int_>(state)) (r.*by_index)(state.cast<int>(), f);
                else (r.*by_coefs)(state.cast<std::vector<cdouble>>(), f);
            } catch (...) {
                std::fclose(f);
                throw;
            }
            std::fclose(f);
        }, py::arg("state"), py::arg("filename"));
    };
    def_state_writer("writeExtendedReciprocalAmplitude", &ResultTB::writeExtendedReciprocalAmplitude,
                     &ResultTB::writeExtendedReciprocalAmplitude);
    def_state_writer("writePhase", &ResultTB::writePhase, &ResultTB::writePhase);
    def_state_writer("writeExtendedPhase", &ResultTB::writeExtendedPhase, &ResultTB::writeExtendedPhase);

    py::class_<ExcitonTB>(m, "ExcitonTB")
        .def(py::init<const SystemConfiguration&, int, int, int, const std::vector<double>&, const std::vector<double>&>(),
             py::arg("config"), py::arg("ncell") = 20, py::arg("nbands") = 1, py::arg("nrmbands") = 0,
             py::arg("parameters") = std::vector<double>{1, 5, 1}, py::arg("Q") = std::vector<double>{0., 0., 0.})
        .def(py::init<const SystemConfiguration&, int, const std::vector<long long>&, const std::vector<double>&,
                      const std::vector<double>&>(),
             py::arg("config"), py::arg("ncell"), py::arg("bands"), py::arg("parameters") = std::vector<double>{1, 5, 1},
             py::arg("Q") = std::vector<double>{0., 0., 0.})
        .def(py::init<const SystemConfiguration&, const ExcitonConfiguration&>())
        .def(py::init<std::shared_ptr<SystemTB>, int, int, int, const std::vector<double>&, const std::vector<double>&>(),
             py::arg("system"), py::arg("ncell") = 20, py::arg("nbands") = 1, py::arg("nrmbands") = 0,
             py::arg("parameters") = std::vector<double>{1, 5, 1}, py::arg("Q") = std::vector<double>{0., 0., 0.})
        .def("setParameters", py::overload_cast<const std::vector<double>&>(&ExcitonTB::setParameters))
        .def("setGauge", &ExcitonTB::setGauge)
        .def("setMode", &ExcitonTB::setMode)
        .def("setReciprocalVectors", &ExcitonTB::setReciprocalVectors)
        .def("setRegularization", &ExcitonTB::setRegularization)
        .def("setUnitCells", &ExcitonTB::setUnitCells)
        .def("setBands", py::overload_cast<const std::vector<long long>&>(&ExcitonTB::setBands))
        .def("setQ", &ExcitonTB::setQ)
        .def("setCutoff", &ExcitonTB::setCutoff)
        .def("setScissor", &ExcitonTB::setScissor)
        .def("setExchange", &ExcitonTB::setExchange)
        .def("setPotential", &ExcitonTB::setPotential)
        .def("setExchangePotential", &ExcitonTB::setExchangePotential)
        .def("setPrecision", &ExcitonTB::setPrecision)
        .def("setMatrixFree", &ExcitonTB::setMatrixFree)
        .def("setDavidsonBlock", &ExcitonTB::setDavidsonBlock)
        .def("setDavidsonTolerance", &ExcitonTB::setDavidsonTolerance)
        .def("setDevice", &ExcitonTB::setDevice)
        .def("brillouinZoneMesh", &ExcitonTB::brillouinZoneMesh)
        .def("initializeHamiltonian", &ExcitonTB::initializeHamiltonian, py::call_guard<py::gil_scoped_release>())
        .def("BShamiltonian", [](ExcitonTB& e) {
            py::gil_scoped_release rel;
            e.BShamiltonian();
        })
        .def("BShamiltonian",
             [](ExcitonTB& e, py::array_t<long long, py::array::c_style | py::array::forcecast> basis) {
                 if (basis.ndim() != 2 || basis.shape(1) != 3) {
                     throw std::invalid_argument("BShamiltonian(basis): basis must be an n x 3 integer array");
                 }
                 std::vector<long long> b(basis.data(), basis.data() + basis.size());
                 py::gil_scoped_release rel;
                 e.BShamiltonian(b);
             })
        .def("diagonalize", &ExcitonTB::diagonalize, py::arg("method") = "diag", py::arg("nstates") = 8,
             py::call_guard<py::gil_scoped_release>())
        .def_property_readonly("system", [](const ExcitonTB& e) { return e.system; })
        .def_readonly("ncell", &ExcitonTB::ncell)
        .def_readonly("nbands", &ExcitonTB::nbands)
        .def_readonly("nrmbands", &ExcitonTB::nrmbands)
        .def_readonly("totalCells", &ExcitonTB::totalCells)
        .def_readonly("excitonbasisdim", &ExcitonTB::excitonbasisdim)
        .def_readonly("eps_m", &ExcitonTB::eps_m)
        .def_readonly("eps_s", &ExcitonTB::eps_s)
        .def_readonly("r0", &ExcitonTB::r0)
        .def_readonly("regularization", &ExcitonTB::regularization)
        .def_readonly("cutoff", &ExcitonTB::cutoff)
        .def_readonly("scissor", &ExcitonTB::scissor)
        .def_readonly("exchange", &ExcitonTB::exchange)
        .def_readonly("mode", &ExcitonTB::mode)
        .def_readonly("gauge", &ExcitonTB::gauge)
        .def_readonly("potential", &ExcitonTB::potential)
        .def_readonly("exchangePotential", &ExcitonTB::exchangePotential)
        .def_readonly("bands", &ExcitonTB::bands)
        .def_readonly("valenceBands", &ExcitonTB::valenceBands)
        .def_readonly("conductionBands", &ExcitonTB::conductionBands)
        .def_readonly("bandList", &ExcitonTB::bandList)
        .def_readonly("Q", &ExcitonTB::Q)
        .def_readonly("bandToIndex", &ExcitonTB::bandToIndex)
        .def_readonly("precision", &ExcitonTB::precision)
        .def_property_readonly("basisStates", [](const ExcitonTB& e) { return irows(e.basisStates, 3); })
        .def_property_readonly("eigvalKStack",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t nb = e.bandList.size();
                                   return fortran(e.eigvalKStack, {nb, e.system->nk});
                               })
        .def_property_readonly("eigvalKQStack",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t nb = e.bandList.size();
                                   return fortran(e.eigvalKQStack, {nb, e.system->nk});
                               })
        .def_property_readonly("eigvecKStack",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t nb = e.bandList.size();
                                   return fortran(e.eigvecKStack, {e.system->basisdim, nb, e.system->nk});
                               })
        .def_property_readonly("eigvecKQStack",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t nb = e.bandList.size();
                                   return fortran(e.eigvecKQStack, {e.system->basisdim, nb, e.system->nk});
                               })
        .def_property_readonly("HBS",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t n = e.hbsDim();
                                   return fortran(e.HBS(), {n, n});
                               })
        .def_property_readonly("HBS32",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t n = e.excitonbasisdim;
                                   return fortran(e.HBS32(), {n, n});
                               })
        .def_property_readonly("ftMotifStack",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t na = e.system->natoms;
                                   std::vector<cdouble> v = e.ftMotifStack();
                                   return fortran(v, {na, na, static_cast<py::ssize_t>(v.size() / (na * na))});
                               })
        .def_property_readonly("ftMotifQ",
                               [](const ExcitonTB& e) {
                                   const py::ssize_t na = e.system->natoms;
                                   return fortran(e.ftMotifQ(), {na, na});
                               })
        .def("cellCoefficients",
             [](const ExcitonTB& e) {
                 std::vector<int> c = e.cellCoefficients();
                 std::vector<long long> v(c.begin(), c.end());
                 return irows(v, e.system->ndim);
             })
        .def("applyBSE",
             [](const ExcitonTB& e, py::array_t<cdouble, py::array::f_style | py::array::forcecast> x,
                const std::string& prec, bool mf) {
                 const int nvec = x.ndim() == 1 ? 1 : static_cast<int>(x.shape(1));
                 std::vector<cdouble> y = e.applyBSE(to_vec(x), nvec, prec, mf);
                 return fortran(y, {static_cast<py::ssize_t>(e.excitonbasisdim), nvec});
             },
             py::arg("x"), py::arg("precision") = "fp64", py::arg("matrixFree") = false)
        .def("hfGradients",
             [](const ExcitonTB& e, py::object eigvec, int nstates, const std::vector<std::string>& params) {
                 std::vector<cdouble> v;
                 if (!eigvec.is_none()) v = to_vec(eigvec.cast<py::array_t<cdouble, py::array::f_style | py::array::forcecast>>());
                 std::vector<double> g = e.hfGradients(v, nstates, params);
                 return fortran(g, {nstates, static_cast<py::ssize_t>(params.size())});
             },
             py::arg("eigvec"), py::arg("nstates"), py::arg("params"))
        .def("hoppingGradients",
             [](const ExcitonTB& e, py::object eigvec, int nstates) {
                 std::vector<cdouble> v;
                 if (!eigvec.is_none()) v = to_vec(eigvec.cast<py::array_t<cdouble, py::array::f_style | py::array::forcecast>>());
                 std::vector<cdouble> G = e.hoppingGradients(v, nstates);
                 const py::ssize_t bd = e.system->basisdim, nR = e.system->ncells;
                 // [state][cell][b][a] in memory -> array indexed [state, cell, a, b]
                 py::array_t<cdouble> out({static_cast<py::ssize_t>(nstates), nR, bd, bd},
                                          {static_cast<py::ssize_t>(nR * bd * bd * sizeof(cdouble)),
                                           static_cast<py::ssize_t>(bd * bd * sizeof(cdouble)),
                                           static_cast<py::ssize_t>(sizeof(cdouble)),
                                           static_cast<py::ssize_t>(bd * sizeof(cdouble))});
                 std::copy(G.begin(), G.end(), out.mutable_data());
                 return out;
             },
             py::arg("eigvec"), py::arg("nstates"))
        .def("timings", &ExcitonTB::timings)
        .def("lastSolveReport", &ExcitonTB::lastSolveReport)
        .def("contextAddress", [](const ExcitonTB& e) { return reinterpret_cast<std::uintptr_t>(e.context()); });

    m.def("detectDegeneracies", &detectDegeneracies, py::arg("eigval"), py::arg("n"), py::arg("precision") = 6);
    // The host Hermitian eigensolver (bands, Rayleigh-Ritz), exposed for testing.
    m.def("hermitianEigh", [](py::array_t<cdouble, py::array::f_style | py::array::forcecast> a) {
        const int n = static_cast<int>(a.shape(0));
        if (a.ndim() != 2 || a.shape(1) != n) throw std::invalid_argument("square matrix expected");
        std::vector<double> w(n);
        std::vector<cdouble> z(static_cast<size_t>(n) * n);
        xatu_linalg::heev(n, a.data(), n, w.data(), z.data());
        return py::make_tuple(py::array_t<double>(n, w.data()), fortran(z, {n, n}));
    });
}
