"""Result observables and writers of the facade (ResultTB: spinX, realSpaceWavefunction,
binding energy, phase and extended writers) on states solved on the GPU, against the
reference's regression values (test/tests.cpp:323-366, 627-656) and the oracle
restatements applied to the facade's own band stacks and states."""
import numpy as np
import pytest

import oracle as O
import xatu_amd as X
from cases import MODELS, oracle_case

pytestmark = pytest.mark.gpu

HBN = MODELS / "hBN.model"
MOS2 = MODELS / "MoS2.model"

SPIN_MOS2 = np.array([[-9.915e-01, -5.000e-01, -4.915e-01], [-6.966e-04, -5.000e-01, 4.993e-01],
                      [8.297e-03, 4.998e-01, -4.915e-01], [9.991e-01, 4.998e-01, 4.993e-01]])


def with_facade_stacks(oc, ex):
    """Give the oracle case the facade's band stacks (same band gauge as its states)."""
    oc.eigvec_k = np.asfortranarray(np.asarray(ex.eigvecKStack).reshape(oc.eigvec_k.shape, order="F"))
    oc.eigvec_kq = np.asfortranarray(np.asarray(ex.eigvecKQStack).reshape(oc.eigvec_kq.shape, order="F"))
    oc.eigval_k = np.asfortranarray(np.asarray(ex.eigvalKStack).reshape(oc.eigval_k.shape, order="F"))
    return oc


def hbn20(nstates=2):
    ex = X.ExcitonTB(X.SystemConfiguration(str(HBN)), 20, 1, 0, [1, 1, 10])
    ex.brillouinZoneMesh(20)
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    return ex, ex.diagonalize("diag", nstates)


def test_spin_reference_values():
    # test/tests.cpp:627-656
    ex = X.ExcitonTB(X.SystemConfiguration(str(MOS2)), 12, 2, 0, [1., 4., 13.55])
    ex.system.reducedBrillouinZoneMesh(12, 2)
    ex.system.shiftBZ([0.6628, -1.1480, 0])
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    res = ex.diagonalize("diag", 4)
    spin = np.array([np.asarray(res.spinX(n)).real for n in range(4)])
    assert np.abs(spin - SPIN_MOS2).max() < 1e-4
    oc = with_facade_stacks(oracle_case(MOS2, 12, 2, 0, (1., 4., 13.55), factor=2, shift=(0.6628, -1.1480, 0)), ex)
    v = np.asarray(res.eigvec)
    for n in range(4):
        assert np.abs(np.asarray(res.spinX(list(v[:, n]))) - O.exciton_spin(oc, v[:, n])).max() < 1e-12


def test_spin_needs_spin_orbitals():
    ex, res = hbn20()
    with pytest.raises(Exception, match="spin orbitals"):
        res.spinX(0)


def test_real_space_wavefunction_reference_hash(tmp_path):
    # test/tests.cpp:323-366
    ex, res = hbn20()
    s = ex.system
    cells = np.asarray(s.truncateSupercell(20, np.linalg.norm(np.asarray(s.bravaisLattice)[0]) * 20))
    v = np.asarray(res.eigvec)
    rswf = np.zeros(len(cells) * 2)
    for n in range(2):
        it = 0
        for cell in cells:
            for a in range(2):
                rswf[it] += res.realSpaceWavefunction(list(v[:, n]), a, 1, list(cell), [0.0, 0.0, 0.0])
                it += 1
    assert abs(O.array2hash(rswf) - 3.4185866144) < 1e-5
    # the writer: hole position, one line per (cell, atom), "#"
    f = tmp_path / "rs.txt"
    res.writeRealspaceAmplitude(0, 1, [0.0, 0.0, 0.0], str(f), 20)
    lines = f.read_text().splitlines()
    assert lines[-1] == "#" and len(lines) == 2 + len(cells) * 2
    tau1 = np.asarray(s.motif)[1, :2]
    assert np.allclose([float(t) for t in lines[0].split()[:2]], tau1, atol=1e-8)
    vals = np.array([float(l.split()[2]) for l in lines[1:-1]])
    ref0 = np.array([res.realSpaceWavefunction(list(v[:, 0]), a, 1, list(c), [0.0, 0.0, 0.0]) for c in cells
                     for a in range(2)])
    assert np.allclose(vals, ref0, atol=1e-10)


def test_real_space_wavefunction_matches_oracle():
    oc = oracle_case(MOS2, 6, 2, 0, (1., 4., 13.55), Q=(0, 0.05, 0))
    ex = X.ExcitonTB(X.SystemConfiguration(str(MOS2)), 6, 2, 0, [1., 4., 13.55], [0, 0.05, 0])
    ex.brillouinZoneMesh(6)
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    res = ex.diagonalize("davidson", 3)
    oc = with_facade_stacks(oc, ex)
    v = np.asarray(res.eigvec)
    rng = np.random.default_rng(4)
    for _ in range(6):
        e, h = rng.integers(0, 3, size=2)
        ecell = rng.integers(-2, 3, size=2) @ np.asarray(ex.system.bravaisLattice)[:2]
        got = res.realSpaceWavefunction(list(v[:, 1]), int(e), int(h), list(ecell), [0.0, 0.0, 0.0])
        ref = O.real_space_wavefunction(oc, v[:, 1], e, h, ecell, np.zeros(3))
        assert got == pytest.approx(ref, rel=1e-12, abs=1e-15)


def test_gap_binding_energy_and_envelope():
    ex, res = hbn20(4)
    # hBN bands: E = +-sqrt(3.625^2 + 2.3^2 |f(k)|^2), gap 7.25 eV at K; the mesh point of
    # the ground-state peak is at or next to K
    gap = res.determineGap()
    assert 7.25 <= gap < 7.6
    k = res.findExcitonPeak(0)
    v0 = np.abs(np.asarray(res.eigvec)[:, 0])
    assert k == int(np.argmax(v0))  # one band pair per k
    assert res.bindingEnergy(0) == pytest.approx(res.eigval[0] - gap, abs=1e-12)
    assert res.bindingEnergy(0, 7.0) == pytest.approx(res.eigval[0] - 7.0, abs=1e-12)
    # at r_e = r_h the envelope transform is sum_k |A_k|^2 = 1 for a normalised state
    assert res.ftExcitonEnvelope(0, [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]) == pytest.approx(1.0, abs=1e-12)
    r = [2.5, 1.0, 0.0]
    kp = np.asarray(ex.system.kpoints)
    ref = np.sum(np.abs(np.asarray(res.eigvec)[:, 0]) ** 2 * np.cos(kp @ np.asarray(r)))
    assert res.ftExcitonEnvelope(0, r, [0.0, 0.0, 0.0]) == pytest.approx(ref, abs=1e-12)
    assert res.boundingBoxBZ() == pytest.approx(max(kp[:, 0].max(), kp[:, 1].max()))


def test_phase_and_extended_writers(tmp_path):
    ex, res = hbn20()
    kp = np.asarray(ex.system.kpoints)
    v0 = np.asarray(res.eigvec)[:, 0]
    f = tmp_path / "phase.txt"
    res.writePhase(0, str(f))
    rows = [l.split() for l in f.read_text().splitlines()[1:-1]]
    got = np.array([[float(x) for x in r] for r in rows])
    assert np.allclose(got[:, :3], kp, atol=1e-8)
    assert np.allclose(got[:, 3], np.abs(v0), atol=1e-8)
    assert np.allclose(got[:, 4], np.angle(v0), atol=1e-8)
    # extended writers: every printed point is a mesh point shifted by a combination of
    # the reciprocal vectors with coefficients in {-1, 0, 1}, inside the box
    box = res.boundingBoxBZ()
    B = np.asarray(ex.system.reciprocalLattice)
    dk = np.linalg.norm(kp[1] - kp[0])
    for name in ("writeExtendedReciprocalAmplitude", "writeExtendedPhase"):
        f = tmp_path / f"{name}.txt"
        getattr(res, name)(0, str(f))
        lines = [l for l in f.read_text().splitlines() if l and l[0] not in "k#"]
        pts = np.array([[float(x) for x in l.split()] for l in lines])
        assert np.all(np.abs(pts[:, 0]) < box) and np.all(np.abs(pts[:, 1]) < box)
        shifts = np.array([(i - 1) * B[0] + (j - 1) * B[1] for j in range(3) for i in range(3)])
        nearest = np.array([np.min(np.linalg.norm(kp[:, None, :] + shifts[None] - p[:3], axis=2)) for p in pts[:50]])
        assert nearest.max() < 1e-6
        if name == "writeExtendedReciprocalAmplitude":
            amp = np.asarray(res.reciprocalAmplitude(0))
            assert np.isclose(pts[:, 3].max(), amp.max(), rtol=1e-6)
        else:
            assert np.isclose(pts[:, 3].max(), np.abs(v0).max() / dk, rtol=1e-6)
    with pytest.raises(Exception, match="one valence and conduction"):
        ex2 = X.ExcitonTB(X.SystemConfiguration(str(MOS2)), 6, 2, 0, [1., 4., 13.55])
        ex2.brillouinZoneMesh(6)
        ex2.initializeHamiltonian()
        ex2.BShamiltonian()
        ex2.diagonalize("davidson", 2).writePhase(0, str(tmp_path / "x.txt"))


def test_write_spin(tmp_path):
    ex = X.ExcitonTB(X.SystemConfiguration(str(MOS2)), 12, 2, 0, [1., 4., 13.55])
    ex.system.reducedBrillouinZoneMesh(12, 2)
    ex.system.shiftBZ([0.6628, -1.1480, 0])
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    res = ex.diagonalize("davidson", 4)
    f = tmp_path / "spin.txt"
    res.writeSpin(4, str(f))
    lines = f.read_text().splitlines()
    assert lines[0].split() == ["n", "St", "Se", "Sh"] and len(lines) == 5
    vals = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    assert np.abs(vals[:, 1:] - SPIN_MOS2).max() < 2e-4
