"""GPU parity at the configured workload sizes.

  C1  hBN two-band model, 30 x 30 mesh, cutoff 20 cells, exchange on, N = 900: lattice sums,
      FP64 matrix and the full dense diagonalisation (all 900 pairs) against the oracle.
  C2  seeded 18-orbital model, 48 x 48, nv = nc = 2, N = 9216: lattice sums and the whole
      FP64 matrix against the oracle; the lowest 16 energies (FP64 Davidson, mixed form with
      either FP32-matrix product) against the eigenvalues of the oracle's matrix.
  C3  same model, 36 x 36, N = 5184: Hellmann-Feynman gradients of the lowest 8 energies with
      respect to r0 and eps_s against the oracle's explicit dH; the hopping gradients against
      central differences along a random Hermitian change of every H(R).
  C4  same model, 96 x 96, N = 36864: lattice sums; 64 sampled columns of the FP64 matrix (unit
      vector products) against the oracle's element rule; the lowest 16 FP64 and mixed energies,
      bounded by FP64 residuals; the convolution-form operator against the resident matrix.
  C5  seeded 36-orbital model, 120 x 120, nv = nc = 3, N = 129600, matrix-free: lattice sums;
      64 sampled columns of the convolution-form operator against the oracle; the lowest 16
      energies of the mixed matrix-free solve, with residuals from the recomputed-tiles operator
      (an independent evaluation of H).
Tolerances: lattice sums 1e-12 relative, matrix entries 1e-10 relative, FP64 energies 1e-6 eV,
mixed energies 0.1 meV, gradients 1e-9 (explicit dH) and 1e-4 relative (differences).
"""
import numpy as np
import pytest

import oracle as O
from cases import MODELS, gpu_context, oracle_case, rel_err

pytestmark = pytest.mark.gpu

R0, EPS_M, EPS_S = 0, 1, 2


def say(msg):
    print(f"[configs] {msg}", flush=True)


def check_fourier(ctx, ex):
    ft, ftq = ctx.get_fourier()
    assert rel_err(ft, ex.ft) < 1e-12
    assert rel_err(ftq, ex.ftq) < 1e-12


def residuals(apply, w, v):
    r = apply(v) - v * w
    return np.linalg.norm(r, axis=0)


def test_c1_full_dense_diagonalisation():
    say("C1 hBN 30x30 cutoff 20, full diag")
    ex = oracle_case(MODELS / "hBN.model", 30, 1, 0, (1, 1, 10), cutoff=20, exchange=True)
    H = ex.bse()
    wr = np.linalg.eigvalsh(H)
    ctx = gpu_context(ex)
    check_fourier(ctx, ex)
    Hg = ctx.build(0, copy_out=True)
    assert rel_err(Hg, H) < 1e-10
    w, v, _ = ctx.solve(method=0, nstates=ex.N)
    ctx.close()
    assert len(w) == 900 and np.abs(w - wr).max() < 1e-9
    assert np.abs(H @ v - v * w).max() < 1e-9


def test_c2_matrix_and_lowest_energies(synthetic18):
    say("C2 seeded 18-orbital 48x48 N=9216: oracle build")
    ex = oracle_case(synthetic18, 48, 2, 0, (1, 1, 40), cutoff=25, exchange=True)
    H = ex.bse()
    ctx = gpu_context(ex)
    check_fourier(ctx, ex)
    say("C2 FP64 build")
    Hg = ctx.build(0, copy_out=True)
    assert rel_err(Hg, H) < 1e-10
    assert np.array_equal(Hg, Hg.conj().T)
    del Hg
    say("C2 oracle eigenvalues")
    wr = np.linalg.eigvalsh(H)[:16]
    w64, v64, rep = ctx.solve(method=1, nstates=16, precision=0, block=32, tol=1e-6)
    assert rep.converged == 1
    assert np.abs(w64 - wr).max() < 1e-6
    assert residuals(lambda x: H @ x, w64, v64).max() < 2e-6
    ctx.build(1)
    for arith in (0, 1):  # split-FP16 and FP32 matrix-core products of the FP32 matrix
        say(f"C2 mixed solve, arithmetic {arith}")
        w, v, rep = ctx.solve(method=1, nstates=16, precision=1, block=32, tol=1e-5, fp32_arith=arith)
        assert rep.converged == 1
        assert np.abs(w - wr).max() < 1e-4, (w - wr) * 1e3
        assert residuals(lambda x: H @ x, w, v).max() < 2e-5
    ctx.close()


def _hermitian_direction(s, rng):
    """Random change of every H(R) with dH(-R) = dH(R)^H (dH(0) Hermitian)."""
    dH = rng.normal(size=s.H.shape) + 1j * rng.normal(size=s.H.shape)
    for r in range(len(s.cells)):
        rp = int(np.argmin(np.linalg.norm(s.cells + s.cells[r], axis=1)))
        if rp == r:
            dH[r] = 0.5 * (dH[r] + dH[r].conj().T)
        elif rp > r:
            dH[rp] = dH[r].conj().T
    return dH * (np.abs(s.H).max() / np.abs(dH).max())


def test_c3_gradients(synthetic18):
    import xatu_amd as X
    say("C3 seeded 18-orbital 36x36 N=5184: oracle build")
    kw = dict(cutoff=25, exchange=True)
    ex = oracle_case(synthetic18, 36, 2, 0, (1, 1, 40), **kw)
    H = ex.bse()
    ns = 8
    ctx = gpu_context(ex)
    check_fourier(ctx, ex)
    ctx.build(0)
    w, psi, _ = ctx.solve(method=0, nstates=ns + 1)
    assert np.abs(H @ psi - psi * w).max() < 1e-9
    assert np.min(np.diff(w)) > 1e-5, w  # non-degenerate: the gradients are defined
    psi = psi[:, :ns]
    say("C3 screening gradients")
    g = ctx.hf_gradients(ns, [R0, EPS_S], eigvec=psi)
    ctx.close()
    for j, kind in ((0, O.POT_KELDYSH_DR0), (1, O.POT_KELDYSH_DEPS)):
        dft = ex.ft_stack(kind)
        dH = ex.bse(ft=dft, ftq=ex.ft_at(ex.Q, kind), eigval_k=np.zeros_like(ex.eigval_k),
                    eigval_kq=np.zeros_like(ex.eigval_kq), scissor=0.0)
        ref = np.real(np.einsum("in,ij,jn->n", psi.conj(), dH, psi))
        del dH
        assert np.abs(g[:, j] - ref).max() <= 1e-9 * np.abs(ref).max(), j
    say("C3 hopping gradients")
    fx = X.ExcitonTB(X.SystemConfiguration(synthetic18), 36, 2, 0, [1.0, 1.0, 40.0])
    fx.setCutoff(25)
    fx.setExchange(True)
    fx.brillouinZoneMesh(36)
    fx.initializeHamiltonian()
    fx.BShamiltonian()
    res = fx.diagonalize("diag", ns)
    assert np.abs(np.asarray(res.eigval)[:ns] - w[:ns]).max() < 1e-9
    G = np.asarray(fx.hoppingGradients(None, ns))  # ns x ncells x norb x norb
    base = O.System(synthetic18)
    dH = _hermitian_direction(base, np.random.default_rng(5))
    analytic = np.real(np.einsum("srab,rab->s", G, dH))
    h = 1e-5

    def energies(sign):
        s = O.System(synthetic18)
        s.H = s.H + sign * h * dH
        s.brillouin_zone_mesh(36)
        e = O.Exciton(s, 36, 2, 0, (1, 1, 40))
        e.cutoff, e.exchange = 25, True
        e.initialize()
        c = gpu_context(e)
        c.build(0)
        we, _, _ = c.solve(method=0, nstates=ns, want_vectors=False)
        c.close()
        return we[:ns]

    fd = (energies(+1) - energies(-1)) / (2 * h)
    say(f"C3 hopping directional derivative: analytic {analytic}, differences {fd}")
    assert np.all(np.abs(fd - analytic) <= 1e-4 * np.abs(analytic).max()), (fd, analytic)


def test_c4_columns_and_energies(synthetic18):
    say("C4 seeded 18-orbital 96x96 N=36864: oracle lattice sums")
    ex = oracle_case(synthetic18, 96, 2, 0, (1, 1, 40), cutoff=25, exchange=True)
    ctx = gpu_context(ex)
    check_fourier(ctx, ex)
    say("C4 FP64 build, sampled columns")
    ctx.build(0)
    rng = np.random.default_rng(4)
    cols = np.sort(rng.choice(ex.N, size=64, replace=False))
    E = np.zeros((ex.N, 64), complex, order="F")
    E[cols, np.arange(64)] = 1.0
    Hc = ctx.matvec(E, precision=0)
    ref = ex.bse_columns(cols)
    assert np.abs(Hc - ref).max() <= 1e-10 * np.abs(ref).max()
    del E, Hc
    say("C4 FP64 Davidson")
    w64, v64, rep = ctx.solve(method=1, nstates=16, precision=0, block=32, tol=1e-5)
    assert rep.converged == 1
    r64 = residuals(lambda x: ctx.matvec(x, precision=0), w64, v64)
    assert r64.max() <= 1e-5
    # the convolution-form operator (no matrix) equals the resident one on the Ritz vectors
    y_dense = ctx.matvec(v64, precision=0)
    assert rel_err(ctx.matvec(v64, precision=0, matrix_free=1), y_dense) < 1e-12
    say("C4 mixed solve")
    ctx.build(1)
    for arith in (0, 1):
        w, v, rep = ctx.solve(method=1, nstates=16, precision=1, block=32, tol=1e-5, fp32_arith=arith)
        assert rep.converged == 1
        # both forms bounded by FP64 residuals of 1e-5 eV; their energies agree within 0.1 meV
        assert residuals(lambda x: ctx.matvec(x, precision=0), w, v).max() <= 1.2e-5
        assert np.abs(w - w64).max() < 1e-4, (w - w64) * 1e3
    ctx.close()


def test_c5_matrix_free(synthetic36):
    say("C5 seeded 36-orbital 120x120 N=129600: oracle lattice sums")
    ex = oracle_case(synthetic36, 120, 3, 0, (1, 1, 40), cutoff=25, exchange=True)
    ctx = gpu_context(ex)
    check_fourier(ctx, ex)
    ctx.build(2)  # matrix-free: exchange factors and the diagonal only
    say("C5 sampled columns of the convolution-form operator")
    rng = np.random.default_rng(5)
    cols = np.sort(rng.choice(ex.N, size=64, replace=False))
    E = np.zeros((ex.N, 64), complex, order="F")
    E[cols, np.arange(64)] = 1.0
    Hc = ctx.matvec(E, precision=0, matrix_free=1)
    ref = ex.bse_columns(cols)
    assert np.abs(Hc - ref).max() <= 1e-10 * np.abs(ref).max()
    del E, Hc
    say("C5 mixed matrix-free solve")
    w, v, rep = ctx.solve(method=1, nstates=16, precision=1, matrix_free=1, block=32, tol=1e-5)
    assert rep.converged == 1
    say("C5 residuals with the recomputed-tiles operator")
    r = residuals(lambda x: ctx.matvec(x, precision=0, matrix_free=2), w, v)
    assert r.max() <= 1.5e-5, r
    ctx.close()
