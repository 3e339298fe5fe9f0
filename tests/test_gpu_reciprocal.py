"""Reciprocal-space mode (ExcitonTB::setMode("reciprocalspace"): src/ExcitonTB.cpp:408-426,
562-621, 895-916) on the device, through the C-ABI and the facade:
  FP64 matrix      <= 1e-10 relative to max |H| against the oracle restatement, exactly Hermitian
  FP32 matrix      per-element error at FP32 rounding; mixed energies <= 1e-4 eV
  products         dense and matrix-free (the convolution form is served by recomputed tiles)
  energies         the reference's regression values (test/tests.cpp:246-279)
"""
import numpy as np
import pytest

import oracle as O
from cases import MODELS, gpu_context, oracle_case, rel_err

pytestmark = pytest.mark.gpu

HBN = MODELS / "hBN.model"
MOS2 = MODELS / "MoS2.model"

CASES = {
    "hbn8_n3": dict(kw=dict(model=HBN, ncell=8, nbands=1, params=(1, 1, 10)), nrec=3),
    # even n_reciprocal: centred offsets -2..1 (asymmetric set of G)
    "hbn9_n4_x": dict(kw=dict(model=HBN, ncell=9, nbands=1, params=(1, 1, 10), exchange=True, scissor=0.2), nrec=4),
    "mos2_q_x": dict(kw=dict(model=MOS2, ncell=6, nbands=2, params=(1., 4., 13.55), Q=(0, 0.07, 0),
                             exchange=True), nrec=3),
    "mos2_odd": dict(kw=dict(model=MOS2, ncell=5, nbands=2, params=(1., 4., 13.55), cutoff=1.2), nrec=5),
}

_cache = {}


def get_case(name):
    if name not in _cache:
        spec = CASES[name]
        print(f"[reciprocal {name}] oracle build", flush=True)
        ex = oracle_case(**spec["kw"])
        _cache[name] = (ex, O.bse_reciprocal(ex, spec["nrec"]), spec["nrec"])
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_reciprocal_fp64_matrix(name):
    ex, H, nrec = get_case(name)
    ctx = gpu_context(ex, n_reciprocal=nrec)
    Hg = ctx.build(0, scissor=ex.scissor, copy_out=True)
    assert rel_err(Hg, H) < 1e-10
    assert np.array_equal(Hg, Hg.conj().T)
    ctx.close()


@pytest.mark.parametrize("name", ["hbn9_n4_x", "mos2_q_x"])
def test_reciprocal_subset_basis(name):
    ex, H, nrec = get_case(name)
    ctx = gpu_context(ex, n_reciprocal=nrec)
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(ex.N, size=min(ex.N, 90), replace=False))
    Hs = ctx.build_basis(ex.basis[idx], scissor=ex.scissor)
    assert rel_err(Hs, H[np.ix_(idx, idx)]) < 1e-10
    ctx.close()


@pytest.mark.parametrize("name", ["hbn9_n4_x", "mos2_q_x", "mos2_odd"])
def test_reciprocal_products_and_solvers(name):
    ex, H, nrec = get_case(name)
    ctx = gpu_context(ex, n_reciprocal=nrec)
    ctx.build(0, scissor=ex.scissor)
    H32 = ctx.build(1, scissor=ex.scissor, copy_out=True)
    assert np.abs(H32.astype(np.complex128) - H).max() < 2e-6 * max(1.0, np.abs(H).max())
    rng = np.random.default_rng(9)
    x = rng.normal(size=(ex.N, 37)) + 1j * rng.normal(size=(ex.N, 37))
    y = H @ x
    assert rel_err(ctx.matvec(x, precision=0), y) < 1e-12
    assert rel_err(ctx.matvec(x, precision=1), y) < 1e-5
    for form in (1, 2):
        assert rel_err(ctx.matvec(x, precision=0, matrix_free=form), y) < 1e-12
        assert rel_err(ctx.matvec(x, precision=1, matrix_free=form), y) < 1e-5
    nstates = 6
    wr = np.linalg.eigvalsh(H)[:nstates]
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=0, tol=1e-7)
    assert rep.converged == 1 and np.abs(w - wr).max() < 1e-6
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=1, tol=1e-6)
    assert rep.converged == 1 and np.abs(w - wr).max() < 1e-4
    ctx.close()
    # matrix-free only: Davidson with FP64 products and the mixed form refined matrix-free
    ctx = gpu_context(ex, n_reciprocal=nrec)
    ctx.build(2, scissor=ex.scissor)
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=1, matrix_free=1, tol=1e-6)
    assert rep.converged == 1 and np.abs(w - wr).max() < 1e-6
    ctx.close()


def test_reciprocal_mode_rejects_lattice_sum_outputs():
    from xatu_amd import capi
    ex, _, nrec = get_case("hbn8_n3")
    ctx = gpu_context(ex, n_reciprocal=nrec)
    with pytest.raises(capi.XatuError, match="reciprocal-space mode has no lattice Fourier sums"):
        ctx.get_fourier()
    ctx.build(0)
    ctx.solve(method=1, nstates=2)
    with pytest.raises(capi.XatuError, match="real-space mode only"):
        ctx.hf_gradients(2, [capi.XATU_PARAM_R0])
    ctx.close()


def test_facade_reciprocal_reference_energies():
    # test/tests.cpp:246-279
    import xatu_amd as X
    ex = X.ExcitonTB(X.SystemConfiguration(str(HBN)), 20, 1, 0, [1, 1, 10])
    ex.setMode("reciprocalspace")
    ex.setReciprocalVectors(5)
    ex.brillouinZoneMesh(20)
    ex.initializeHamiltonian()
    ex.BShamiltonian()
    res = ex.diagonalize("diag", 3)
    levels = O.detect_degeneracies(np.asarray(res.eigval), 3, 6)
    for (e, d), (ee, dd) in zip(levels, [(6.234291, 1), (6.236636, 1), (6.731819, 1)]):
        assert abs(e - ee) < 1e-4 and d == dd
    oc = oracle_case(HBN, 20, 1, 0, (1, 1, 10))
    wr = np.linalg.eigvalsh(O.bse_reciprocal(oc, 5))[:3]
    assert np.abs(np.asarray(res.eigval)[:3] - wr).max() < 1e-8
