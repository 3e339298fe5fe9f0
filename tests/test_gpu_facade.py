"""The reference's regression suite (test/tests.cpp) run end to end through the
C++ ExcitonTB facade: host bands -> C-ABI -> HIP lattice sums, assembly and
diagonalisation on the GPU."""
import numpy as np
import pytest

import oracle as O
import xatu_amd as X
from cases import MODELS, oracle_case, rel_err

pytestmark = pytest.mark.gpu

HBN = str(MODELS / "hBN.model")
HBN_SPIN = str(MODELS / "hBN_spinful.model")
MOS2 = str(MODELS / "MoS2.model")


def levels(eigval, n):
    return X.detectDegeneracies(list(eigval), n, 6)


def check_levels(got, expected, tol):
    for (e, d), (ee, dd) in zip(got, expected):
        assert abs(e - ee) < tol, (got, expected)
        assert d == dd, (got, expected)


def hbn(method, nstates=3):
    ex = X.ExcitonTB(X.SystemConfiguration(HBN), 20, 1, 0, [1, 1, 10])
    ex.brillouinZoneMesh(20)
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    return ex, ex.diagonalize(method, nstates)


@pytest.mark.parametrize("method", ["diag", "davidson", "sparse"])
def test_hbn_energies(method):
    # test/tests.cpp:150-244
    ex, res = hbn(method)
    check_levels(levels(res.eigval, 3), [(5.335690, 2), (6.074062, 1)], 1e-4)


def test_hbn_matrix_matches_oracle():
    ex, _ = hbn("diag")
    ref = oracle_case(HBN, 20, 1, 0, (1, 1, 10))
    assert np.allclose(ex.eigvecKStack, ref.eigvec_k, atol=1e-12, rtol=0)
    assert rel_err(ex.HBS, ref.bse()) < 1e-10
    assert rel_err(ex.ftMotifStack, ref.ft) < 1e-12


def test_hbn_kwf_hash():
    # test/tests.cpp:281-321
    ex, res = hbn("diag", 2)
    kwf = sum(np.asarray(res.reciprocalAmplitude(n)) for n in range(2)) + 0j
    assert abs(O.array2hash(kwf) - 1.086145105) < 1e-5


def test_hbn_spinful():
    # test/tests.cpp:401-431
    ex = X.ExcitonTB(X.SystemConfiguration(HBN_SPIN), 20, 2, 0, [1, 1, 10])
    ex.brillouinZoneMesh(20)
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    res = ex.diagonalize("diag", 12)
    check_levels(levels(res.eigval, 12), [(5.335690, 8), (6.074062, 4)], 1e-4)


def mos2(Q=(0, 0, 0), exchange=False, factor=1, shift=None, method="diag", nstates=4):
    ex = X.ExcitonTB(X.SystemConfiguration(MOS2), 12, 2, 0, [1., 4., 13.55], list(Q))
    ex.setExchange(exchange)
    if factor > 1:
        ex.system.reducedBrillouinZoneMesh(12, factor)
        ex.system.shiftBZ(list(shift))
    else:
        ex.brillouinZoneMesh(12)
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    return ex, ex.diagonalize(method, nstates)


def test_mos2_energies():
    # test/tests.cpp:547-577
    _, res = mos2()
    check_levels(levels(res.eigval, 2), [(1.768783, 2)], 1e-4)


@pytest.mark.parametrize("q,expected", [(-0.1, 1.810464), (-0.05, 1.780106), (0.0, 1.768783), (0.05, 1.780106),
                                        (0.1, 1.810464)])
def test_mos2_exciton_bands(q, expected):
    # test/tests.cpp:660-702
    _, res = mos2(Q=(0, q, 0))
    check_levels(levels(res.eigval, 2)[:1], [(expected, 2)], 1e-5)


def test_mos2_exchange():
    # test/tests.cpp:704-737
    _, res = mos2(Q=(0, 0.1, 0), exchange=True)
    check_levels(levels(res.eigval, 4), [(1.810464, 2), (1.825740, 1), (1.855544, 1)], 1e-5)


def test_mos2_reduced_bz():
    # test/tests.cpp:739-774
    _, res = mos2(factor=2, shift=(0.6628, -1.1480, 0))
    check_levels(levels(res.eigval, 4), [(1.795378, 1), (1.809810, 1), (1.935880, 1), (1.950567, 1)], 1e-4)


def test_mos2_kwf_hash():
    # test/tests.cpp:579-619
    ex, res = mos2()
    kwf = sum(np.asarray(res.reciprocalAmplitude(n)) for n in range(2)) + 0j
    assert abs(O.array2hash(kwf) - 1.1814790902) < 1e-5


def test_reciprocal_amplitude_file_format(tmp_path):
    """writeReciprocalAmplitude (Result.tpp:163-177): header, kx ky kz Prob. rows, '#'."""
    ex, res = hbn("diag", 2)
    f = tmp_path / "kwf.txt"
    res.writeReciprocalAmplitude(0, str(f))
    lines = f.read_text().splitlines()
    assert lines[0] == "kx\tky\tkz\tProb." and lines[-1] == "#"
    vals = np.array([[float(x) for x in ln.split("\t")] for ln in lines[1:-1]])
    kp = np.asarray(ex.system.kpoints)
    assert vals.shape == (len(kp), 4)
    assert np.allclose(vals[:, :3], kp, atol=1e-8)
    assert np.allclose(vals[:, 3], np.asarray(res.reciprocalAmplitude(0)), atol=1e-8)


def test_mos2_davidson_fp32_mixed():
    ex = X.ExcitonTB(X.SystemConfiguration(MOS2), 12, 2, 0, [1., 4., 13.55], [0, 0.1, 0])
    ex.setExchange(True)
    ex.brillouinZoneMesh(12)
    ex.initializeHamiltonian()
    ex.BShamiltonian()              # FP64 (kept for the refinement)
    ex.setPrecision("fp32")
    ex.BShamiltonian()              # FP32
    res = ex.diagonalize("davidson", 4)
    check_levels(levels(res.eigval, 4), [(1.810464, 2), (1.825740, 1), (1.855544, 1)], 1e-5)


def test_result_observables(tmp_path):
    ex, res = hbn("diag", 3)
    k = res.kineticEnergy(0)
    p = res.potentialEnergy(0)
    assert k + p == pytest.approx(res.eigval[0], abs=1e-12)
    assert k > res.eigval[0]  # bound state: negative potential energy
    f = tmp_path / "hbn.eigval"
    res.writeEigenvalues(str(f), 3)
    toks = f.read_text().split()
    assert int(toks[0]) == 400 and float(toks[1]) == pytest.approx(5.335690, abs=1e-6)


@pytest.mark.parametrize("method", ["diag", "davidson"])
def test_atomic_gauge(method):
    # ExcitonTB::setGauge("atomic") (src/ExcitonTB.cpp:257-262, 877-884): the BSE built from
    # the atomic-gauge coefficients, against the oracle in the same gauge; the exposed band
    # stacks stay in the lattice gauge
    ex = X.ExcitonTB(X.SystemConfiguration(MOS2), 6, 2, 0, [1., 4., 13.55], [0, 0.05, 0])
    ex.setExchange(True)
    ex.setGauge("atomic")
    ex.brillouinZoneMesh(6)
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    ref = oracle_case(MOS2, 6, 2, 0, (1., 4., 13.55), exchange=True, Q=(0, 0.05, 0))
    # MoS2 has degenerate bands: the oracle takes the facade's (lattice-gauge) stacks, whose
    # phases inside degenerate subspaces are not pinned
    stacks = dict(eigvec_k=np.asarray(ex.eigvecKStack), eigvec_kq=np.asarray(ex.eigvecKQStack))
    H_lat = ref.bse(**stacks)
    ref.gauge = "atomic"
    H = ref.bse(**stacks)
    assert np.abs(H - H_lat).max() > 1e-4
    assert rel_err(ex.HBS, H) < 1e-10
    res = ex.diagonalize(method, 6)
    wr = np.linalg.eigvalsh(H)[:6]
    assert np.abs(np.asarray(res.eigval)[:6] - wr).max() < 1e-6


def test_subset_basis_hamiltonian():
    # ExcitonTB::BShamiltonian(const arma::imat& basis) (src/ExcitonTB.cpp:840-921): the BSE over
    # a subset of the basis rows {valence band, conduction band, k}, diagonalised through the
    # facade; then the full basis again
    ex = X.ExcitonTB(X.SystemConfiguration(MOS2), 6, 2, 0, [1., 4., 13.55])
    ex.setExchange(True)
    ex.brillouinZoneMesh(6)
    ex.initializeHamiltonian()
    ref = oracle_case(MOS2, 6, 2, 0, (1., 4., 13.55), exchange=True)
    # the facade's stacks (phases inside MoS2's degenerate band pairs are not pinned)
    H = ref.bse(eigvec_k=np.asarray(ex.eigvecKStack), eigvec_kq=np.asarray(ex.eigvecKQStack))
    rng = np.random.default_rng(13)
    idx = np.sort(rng.choice(ex.excitonbasisdim, size=61, replace=False))
    basis = np.asarray(ex.basisStates)[idx]
    ex.BShamiltonian(basis)
    Hs = H[np.ix_(idx, idx)]
    assert np.asarray(ex.HBS).shape == (61, 61)
    assert rel_err(ex.HBS, Hs) < 1e-10
    wr = np.linalg.eigvalsh(Hs)
    res = ex.diagonalize("diag", 61)
    assert np.abs(np.asarray(res.eigval) - wr).max() < 1e-9
    res = ex.diagonalize("davidson", 4)
    assert np.abs(np.asarray(res.eigval)[:4] - wr[:4]).max() < 1e-6
    with pytest.raises(Exception, match="not among the exciton's bands"):
        ex.BShamiltonian(np.array([[0, 1, 0]]))
    ex.BShamiltonian()
    assert rel_err(ex.HBS, H) < 1e-10
    res = ex.diagonalize("davidson", 4)
    assert np.abs(np.asarray(res.eigval)[:4] - np.linalg.eigvalsh(H)[:4]).max() < 1e-6
