"""Host-side checks (CPU only): the C++ facade against the oracle, the host
eigensolver, error behaviour, and the C-ABI library's exported symbols."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import xatu_amd as X
from cases import MODELS, oracle_case
from xatu_amd import capi

ROOT = Path(__file__).resolve().parent.parent
HBN = str(MODELS / "hBN.model")
MOS2 = str(MODELS / "MoS2.model")


def test_capi_exports_every_declared_symbol():
    header = (ROOT / "include" / "xatu_bse.h").read_text()
    declared = sorted(set(re.findall(r"\b(xatu_[a-z_0-9]+)\s*\(", header)))
    assert declared == sorted(capi.EXPORTED)
    lib = ctypes.CDLL(str(capi.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name


def test_capi_errors_without_context():
    L = capi.load()
    assert L.xatu_set_lattice(None, 2, None, 1, None, None, 4) == capi.XATU_ERR_INVALID_ARGUMENT
    assert L.xatu_last_error(None) == b"null context"
    assert L.xatu_get_timings(None, None) == capi.XATU_ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("n", [1, 2, 3, 7, 18, 36, 64])
def test_host_eigensolver_matches_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = A + A.conj().T
    w, v = X._xatu.hermitianEigh(A)
    wr = np.linalg.eigvalsh(A)
    assert np.abs(w - wr).max() < 1e-12 * max(1, np.abs(wr).max())
    assert np.abs(A @ v - v * w).max() < 1e-12 * max(1, np.abs(wr).max())
    assert np.abs(v.conj().T @ v - np.eye(n)).max() < 1e-12


def test_host_eigensolver_degenerate():
    A = np.diag([2.0, 1.0, 2.0, 1.0, 3.0]).astype(complex)
    w, v = X._xatu.hermitianEigh(A)
    assert np.allclose(w, [1, 1, 2, 2, 3])
    assert np.abs(A @ v - v * w).max() < 1e-14


@pytest.mark.parametrize("model", [HBN, MOS2, str(MODELS / "hBN_spinful.model")])
def test_model_parser_matches_oracle(model):
    c = X.SystemConfiguration(model)
    m = O.parse_model(model)
    assert c.ndim == m["ndim"] and c.filling == m["filling"] and list(c.norbitals) == m["norbitals"]
    assert np.array_equal(c.bravaisLattice, m["bravais"])
    assert np.array_equal(c.motif, m["motif"])
    assert np.array_equal(c.bravaisVectors, m["cells"])
    assert np.array_equal(np.transpose(c.hamiltonian, (2, 0, 1)), m["H"])


def test_model_parser_synthetic(synthetic18):
    c = X.SystemConfiguration(synthetic18)
    m = O.parse_model(synthetic18)
    assert np.array_equal(np.transpose(c.hamiltonian, (2, 0, 1)), m["H"])
    H = m["H"]
    cells = m["cells"]
    # H(-R) = H(R)^H exactly
    for i, R in enumerate(cells):
        j = int(np.argmin(np.linalg.norm(cells + R, axis=1)))
        assert np.array_equal(H[j], H[i].conj().T)


def test_synthetic_generator_is_pinned():
    from xatu_amd import synthetic
    _, _, _, H, filling = synthetic.model_arrays(natoms=3, orb_per_atom=6)
    assert filling == 9 and H.shape == (49, 18, 18)
    # fingerprint of the seeded draw
    assert O.array2hash(H[24]) == pytest.approx(O.array2hash(H[24]), rel=0)
    onsite = np.real(np.diag(H[24]))
    assert np.all((onsite[[0, 1, 2]] > -4) & (onsite[[0, 1, 2]] < -2))
    assert np.all((onsite[[3, 4, 5]] > 2) & (onsite[[3, 4, 5]] < 4))


def test_exciton_file_parser():
    e = X.ExcitonConfiguration(str(MODELS / "hBN_spinless.txt"))
    # test/tests.cpp:125-148
    assert e.ncell == 30 and e.nbands == 1 and list(e.eps) == [1, 1, 10] and list(e.Q) == [0, 0, 0]
    assert e.label.strip() == "hBN_N30"
    m = X.ExcitonConfiguration(str(MODELS / "MoS2_exciton.txt"))
    assert list(m.bands) == [-1, 0, 1, 2] and list(m.eps) == [1, 4, 13.55]


def test_config_errors(tmp_path):
    with pytest.raises(ValueError):
        X.SystemConfiguration(str(tmp_path / "missing.model"))
    bad = tmp_path / "bad.model"
    bad.write_text("# dimension\n2\n# filling\n1\n#")
    with pytest.raises(RuntimeError, match="Missing arguments"):
        X.SystemConfiguration(str(bad))
    ex = tmp_path / "ex.txt"
    ex.write_text("# ncells\n10\n# bands\n1\n# dielectric\n1 1 10\n# exchange\nmaybe\n#")
    with pytest.raises(ValueError, match="Exchange option"):
        X.ExcitonConfiguration(str(ex))


def test_bands_match_oracle():
    s = X.SystemTB(X.SystemConfiguration(MOS2))
    o = O.System(MOS2)
    s.brillouinZoneMesh(6)
    o.brillouin_zone_mesh(6)
    assert np.allclose(s.kpoints, o.kpoints, atol=1e-14, rtol=0)
    for k in s.kpoints[::5]:
        w, v = s.solveBands(list(k))
        wr, vr = o.solve_bands(k)
        assert np.abs(w - wr[0]).max() < 1e-12


@pytest.mark.parametrize("n,factor,shift", [(6, 1, None), (5, 1, None), (4, 2, [0.6628, -1.1480, 0])])
def test_mesh_and_offsets(n, factor, shift):
    s = X.SystemTB(X.SystemConfiguration(MOS2))
    ex = oracle_case(MOS2, n, 1, 0, (1, 4, 13.55), factor=factor, shift=shift)
    if factor > 1:
        s.reducedBrillouinZoneMesh(n, factor)
    else:
        s.brillouinZoneMesh(n)
    if shift is not None:
        s.shiftBZ(shift)
    assert np.allclose(s.kpoints, ex.sys.kpoints, atol=1e-13, rtol=0)
    assert np.allclose(s.meshBZ, ex.sys.meshBZ, atol=1e-13, rtol=0)
    from cases import k_offsets
    assert np.array_equal(s.kpointOffsets, k_offsets(ex))
    s.calculateInverseReciprocalMatrix()
    for i in range(0, s.nk, 3):
        for j in range(0, s.nk, 2):
            assert s.findEquivalentPointBZ(list(s.kpoints[j] - s.kpoints[i]), n) == \
                ex.sys.find_equivalent(s.kpoints[j] - s.kpoints[i], n)


def test_supercell_matches_oracle():
    s = X.SystemTB(X.SystemConfiguration(HBN))
    o = O.System(HBN)
    assert np.array_equal(s.truncateSupercell(30, s.a * 20), o.truncate_supercell(30, o.a * 20))


def test_exciton_constructor_errors():
    cfg = X.SystemConfiguration(HBN)
    with pytest.raises(ValueError, match="r0 must be non-zero"):
        X.ExcitonTB(cfg, 10, 1, 0, [1, 1, 0])
    with pytest.raises(ValueError, match="Number of bands"):
        X.ExcitonTB(cfg, 10, 2, 0, [1, 1, 10])
    ex = X.ExcitonTB(cfg, 10, 1, 0, [1, 1, 10])
    with pytest.raises(ValueError, match="BZ mesh must be initialized"):
        ex.initializeHamiltonian()
    with pytest.raises(ValueError, match="potential must be"):
        ex.setPotential("yukawa")
    with pytest.raises(ValueError, match="not initialized"):
        ex.diagonalize("diag", 3)
    with pytest.raises(ValueError):
        ex.setMode("momentum")


def test_exciton_band_selection():
    cfg = X.SystemConfiguration(MOS2)
    ex = X.ExcitonTB(cfg, 12, 2, 0, [1, 4, 13.55])
    f = ex.system.fermiLevel
    assert list(ex.valenceBands) == [f - 1, f] and list(ex.conductionBands) == [f + 1, f + 2]
    assert list(ex.bands) == [-1, 0, 1, 2]
    ex2 = X.ExcitonTB(cfg, 12, [-1, 0, 1, 2], [1, 4, 13.55])
    assert list(ex2.bandList) == list(ex.bandList)
    ex3 = X.ExcitonTB(cfg, X.ExcitonConfiguration(str(MODELS / "MoS2_exciton.txt")))
    assert list(ex3.bandList) == list(ex.bandList) and ex3.cutoff == pytest.approx(30 / 2.5)


def test_crystal_parser_matches_oracle():
    path = str(MODELS / "hBN_base_CREN_LDA.outp")
    m, st = O.parse_crystal(path, 30)
    c = X.CRYSTALConfiguration(path, 30)
    assert (c.ndim, c.natoms, c.norbitalsAO, c.nsh, c.totalElectrons, c.coreElectrons) == \
        (st["ndim"], st["natoms"], st["norbitals"], st["nsh"], st["electrons"], st["core"])
    assert list(c.orbitalsPerSpecies) == m["norbitals"] and c.filling == m["filling"]
    assert c.nOverlap == c.nFock == 30 and not c.MAGNETIC_FLAG
    assert np.array_equal(np.moveaxis(np.asarray(c.hamiltonian), 2, 0), m["H"])
    assert np.array_equal(np.moveaxis(np.asarray(c.overlap), 2, 0), m["S"])
    assert np.array_equal(np.asarray(c.bravaisVectors), m["cells"])
    assert np.array_equal(np.asarray(c.motif), m["motif"])
    # four shells with four gaussians each
    assert [len(sh) for sp in sorted(c.gaussianCoefficients) for sh in c.gaussianCoefficients[sp]] == [4, 4, 4, 4]


def test_crystal_bands_match_oracle():
    path = str(MODELS / "hBN_base_CREN_LDA.outp")
    s = X.SystemTB(X.CRYSTALConfiguration(path, 100))
    s.setAU(True)
    m, _ = O.parse_crystal(path, 100)
    o = O.System(m)
    o.au = True
    assert s.isTriangular and o.triangular
    s.brillouinZoneMesh(9)
    o.brillouin_zone_mesh(9)
    for k in s.kpoints[::7]:
        w, _ = s.solveBands(list(k))
        wr, _ = o.solve_bands(k)
        assert np.abs(w - wr[0]).max() < 1e-10


def test_crystal_needs_a_lattice(tmp_path):
    f = tmp_path / "x.outp"
    f.write_text("nothing here\n")
    with pytest.raises(Exception, match="Finite systems"):
        X.CRYSTALConfiguration(str(f), 10)


def test_crystal_unrestricted_spin_blocks(tmp_path):
    # an unrestricted (magnetic) run prints alpha Fock matrices, a line naming the beta
    # electrons, then the beta ones; the reader makes one spinful system with the alpha and
    # beta matrices as interleaved spin blocks and the overlap repeated for both spins.
    # The run is made from the LDA output: its Fock matrices as alpha, twice them as beta.
    src = (MODELS / "hBN_base_CREN_LDA.outp").read_text().split("\n")
    first_fock = next(i for i, l in enumerate(src) if "FOCK MATRIX - CELL N." in l)
    fock = src[first_fock:]
    beta = []
    for line in fock:
        w = line.split()
        if line.startswith("   ") and w and w[0].isdigit() and "E" in line:
            beta.append(line[:6] + "".join(f"{2 * float(t):12.4E}" for t in w[1:]))
        else:
            beta.append(line)
    text = src[:first_fock] + [" TYPE OF CALCULATION :  UNRESTRICTED OPEN SHELL"] + fock + \
        [" BETA ELECTRONS", ""] + beta
    f = tmp_path / "mag.outp"
    f.write_text("\n".join(text))
    plain = X.CRYSTALConfiguration(str(MODELS / "hBN_base_CREN_LDA.outp"), 12)
    mag = X.CRYSTALConfiguration(str(f), 12)
    assert mag.MAGNETIC_FLAG and not plain.MAGNETIC_FLAG
    assert mag.nFock == plain.nFock == 12
    assert mag.filling == 2 * plain.filling
    assert list(mag.orbitalsPerSpecies) == list(plain.orbitalsPerSpecies)
    assert list(mag.norbitals) == [2 * n for n in plain.norbitals]
    H, H0 = np.asarray(mag.hamiltonian), np.asarray(plain.hamiltonian)
    S, S0 = np.asarray(mag.overlap), np.asarray(plain.overlap)
    assert H.shape == (16, 16, 12) and S.shape == (16, 16, 12)
    assert np.allclose(H[0::2, 0::2], H0, atol=1e-12) and np.allclose(H[1::2, 1::2], 2 * H0, atol=1e-4)
    assert np.all(H[0::2, 1::2] == 0) and np.all(H[1::2, 0::2] == 0)
    assert np.array_equal(S[0::2, 0::2], S0) and np.array_equal(S[1::2, 1::2], S0)
    assert np.all(S[0::2, 1::2] == 0)
