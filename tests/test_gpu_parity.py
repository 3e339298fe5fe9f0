"""GPU parity tests through the C-ABI (include/xatu_bse.h).

Each case builds the oracle pipeline on the CPU, hands the oracle's own band
stacks to the C-ABI, and compares:
  lattice sums   <= 1e-12 relative to max |V(q)|
  FP64 BSE matrix <= 1e-10 relative to max |H| and exactly Hermitian
  FP64 energies  <= 1e-6 eV (dense and Davidson)
  FP32 matrix     per-block error reported, energies (FP32 products + FP64
                  refinement) <= 1e-4 eV (0.1 meV)
"""
import numpy as np
import pytest

from cases import MODELS, gpu_context, oracle_case, rel_err
from xatu_amd import capi

pytestmark = pytest.mark.gpu

HBN = MODELS / "hBN.model"
MOS2 = MODELS / "MoS2.model"

CASES = {
    "hbn20": dict(model=HBN, ncell=20, nbands=1, params=(1, 1, 10)),
    "hbn30_cut20_x": dict(model=HBN, ncell=30, nbands=1, params=(1, 1, 10), cutoff=20, exchange=True),
    "hbn_coulomb": dict(model=HBN, ncell=9, nbands=1, params=(1, 1, 10), potential="coulomb", exchange=True,
                        exchange_potential="coulomb", scissor=0.5),
    "mos2_q_x": dict(model=MOS2, ncell=8, nbands=2, params=(1., 4., 13.55), Q=(0, 0.1, 0), exchange=True),
    "mos2_reduced": dict(model=MOS2, ncell=6, nbands=2, params=(1., 4., 13.55), factor=2,
                         shift=(0.6628, -1.1480, 0)),
    "mos2_odd": dict(model=MOS2, ncell=7, nbands=2, params=(1., 4., 13.55)),
    "random18": dict(model="synthetic18", ncell=12, nbands=2, params=(1, 1, 40), cutoff=25, exchange=True),
    # the shape of the 120 x 120 workload (4 atoms x 9 orbitals, nv = nc = 3) on a small mesh
    "random36": dict(model="synthetic36", ncell=6, nbands=3, params=(1, 1, 40), cutoff=25, exchange=True,
                     Q=(0.0, 0.03, 0.0)),
}

_cache = {}


def get_case(name, request):
    if name not in _cache:
        kw = dict(CASES[name])
        if kw["model"] in ("synthetic18", "synthetic36"):
            kw["model"] = request.getfixturevalue(kw["model"])
        print(f"[case {name}] oracle build", flush=True)
        ex = oracle_case(**kw)
        H = ex.bse()
        _cache[name] = (ex, H)
    return _cache[name]


@pytest.mark.parametrize("name", list(CASES))
def test_fourier_sums(name, request):
    ex, _ = get_case(name, request)
    ctx = gpu_context(ex)
    cells = ctx.get_cells()
    assert len(cells) == len(ex.cells)
    cart = cells @ ex.sys.bravais[:ex.sys.ndim]
    assert np.abs(cart - ex.cells).max() < 1e-12
    ft, ftq = ctx.get_fourier()
    assert rel_err(ft, ex.ft) < 1e-12
    if ex.exchange:
        assert rel_err(ftq, ex.ftq) < 1e-12
    ctx.close()


@pytest.mark.parametrize("name", list(CASES))
def test_bse_fp64_matrix(name, request):
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    Hg = ctx.build(0, scissor=ex.scissor, copy_out=True)
    assert rel_err(Hg, H) < 1e-10
    assert np.array_equal(Hg, Hg.conj().T)
    assert np.all(np.imag(np.diag(Hg)) == 0)
    ctx.close()


@pytest.mark.parametrize("name", ["hbn20", "mos2_q_x", "random18"])
def test_bse_subset_basis(name, request):
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(ex.N, size=min(ex.N, 150), replace=False))
    Hs = ctx.build_basis(ex.basis[idx], scissor=ex.scissor)
    assert rel_err(Hs, H[np.ix_(idx, idx)]) < 1e-10
    ctx.close()


@pytest.mark.parametrize("name", ["hbn20", "mos2_q_x", "mos2_reduced", "random18"])
def test_dense_diag(name, request):
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    ctx.build(0, scissor=ex.scissor)
    w, v, _ = ctx.solve(method=0, nstates=ex.N)
    wr = np.linalg.eigvalsh(H)
    assert np.abs(w - wr).max() < 1e-9
    r = H @ v[:, :8] - v[:, :8] * w[:8]
    assert np.abs(r).max() < 1e-9
    ctx.close()


@pytest.mark.parametrize("name,nstates", [("hbn20", 3), ("hbn30_cut20_x", 8), ("mos2_q_x", 4),
                                          ("random18", 16)])
def test_davidson_fp64(name, nstates, request):
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    ctx.build(0, scissor=ex.scissor)
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=0, tol=1e-7)
    wr = np.linalg.eigvalsh(H)[:nstates]
    assert rep.converged == 1
    assert np.abs(w - wr).max() < 1e-6
    assert np.abs(H @ v - v * w).max() < 1e-6
    ctx.close()


# The Rayleigh-Ritz step runs on the device for blocks up to 32 and bases up to 128, on the
# host otherwise; both, at several basis sizes (restarts), against the oracle's eigenvalues.
@pytest.mark.parametrize("nstates,block,max_basis", [(16, 32, 96), (20, 40, 120), (6, 12, 128), (16, 32, 160)])
def test_davidson_rayleigh_ritz_paths(nstates, block, max_basis, request):
    ex, H = get_case("random18", request)
    ctx = gpu_context(ex)
    ctx.build(0, scissor=ex.scissor)
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=0, block=block, max_basis=max_basis, tol=1e-7)
    wr = np.linalg.eigvalsh(H)[:nstates]
    assert rep.converged == 1
    assert np.abs(w - wr).max() < 1e-6
    assert np.abs(H @ v - v * w).max() < 1e-6
    assert np.abs(v.conj().T @ v - np.eye(nstates)).max() < 1e-8
    ctx.close()


def test_davidson_workspace_reuse(request):
    """Solves of different shapes and forms on one context: the solver workspaces and the
    last solve's vectors are kept between solves and only grow (a larger basis regrows
    them, on the host Rayleigh-Ritz path past 128), so every solve must still match the
    dense eigenvalues, and a new interaction must still invalidate the kept vectors."""
    ex, H = get_case("random18", request)
    wr = np.linalg.eigvalsh(H)
    ctx = gpu_context(ex)
    ctx.build(0, scissor=ex.scissor)
    ctx.build(1, scissor=ex.scissor)
    for nstates, block, max_basis, prec in ((16, 32, 96, 0), (16, 32, 0, 1), (16, 32, 160, 0), (6, 12, 128, 0),
                                            (8, 16, 0, 1), (16, 32, 96, 0)):
        w, v, rep = ctx.solve(method=1, nstates=nstates, precision=prec, block=block, max_basis=max_basis,
                              tol=1e-7 if prec == 0 else 1e-6)
        assert rep.converged == 1, (nstates, block, max_basis, prec)
        assert np.abs(w - wr[:nstates]).max() < (1e-6 if prec == 0 else 1e-4), (nstates, block, max_basis, prec)
        assert np.abs(H @ v - v * w).max() < 1e-5
    # the last solve's vectors serve the gradients without an explicit eigvec ...
    assert np.array_equal(ctx.hf_gradients(4, [0]), ctx.hf_gradients(4, [0], eigvec=v[:, :4]))
    ctx.close()
    # ... until the interaction changes: new lattice sums invalidate them
    ctx = gpu_context(ex)
    ctx.build(0, scissor=ex.scissor)
    ctx.solve(method=1, nstates=4, precision=0, tol=1e-7)
    ctx.fourier_sums(ex.eps_m, ex.eps_s, ex.r0, ex.cutoff, ex.ncell, ex.total_cells,
                     regularization=ex.regularization, exchange=ex.exchange, Q=tuple(ex.Q))
    with pytest.raises(capi.XatuError, match="no eigenvectors"):
        ctx.hf_gradients(4, [0])
    ctx.close()


@pytest.mark.parametrize("arith", [0, 1])  # XATU_ARITH_SPLIT16 (FP16 matrix cores, split operands), XATU_ARITH_FP32
@pytest.mark.parametrize("name,nstates", [("hbn30_cut20_x", 8), ("mos2_q_x", 4), ("random18", 16), ("random36", 8)])
def test_fp32_build_and_mixed_solve(name, nstates, arith, request):
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    ctx.build(0, scissor=ex.scissor)
    H32 = ctx.build(1, scissor=ex.scissor, copy_out=True)
    # FP32 assembly: per-element error at FP32 rounding of O(1 eV) entries
    err = np.abs(H32.astype(np.complex128) - H)
    assert err.max() < 2e-6 * max(1.0, np.abs(H).max())
    blk, mx = ctx.compare_builds(4)
    assert mx == pytest.approx(err.max(), rel=1e-6, abs=1e-12)
    print(f"[{name}] FP32 block errors: max {mx:.3e}, median {np.median(blk):.3e}")
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=1, tol=1e-6, fp32_arith=arith)
    wr = np.linalg.eigvalsh(H)[:nstates]
    print(f"[{name}] mixed (arith {arith}): per-eigenvalue error (meV): {np.round((w - wr) * 1e3, 6)}")
    assert np.abs(w - wr).max() < 1e-4
    assert rep.refine_iterations >= 1
    ctx.close()


# N = 400, 576, 144 (a single row block of the split product) and 196
@pytest.mark.parametrize("name", ["hbn20", "random18", "mos2_reduced", "mos2_odd"])
def test_matvec(name, request):
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    ctx.build(0, scissor=ex.scissor)
    ctx.build(1, scissor=ex.scissor)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(ex.N, 5)) + 1j * rng.normal(size=(ex.N, 5))
    y = H @ x
    assert rel_err(ctx.matvec(x, precision=0), y) < 1e-12
    assert rel_err(ctx.matvec(x, precision=1), y) < 1e-5
    ctx.close()


@pytest.mark.parametrize("name", ["hbn20", "random36"])
def test_split_product_scaling(name, request):
    # per-column power-of-two scales: the relative error does not depend on the magnitude
    # of a column, zero columns stay zero, and fewer than 32 vectors are handled
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    ctx.build(1, scissor=ex.scissor)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(ex.N, 7)) + 1j * rng.normal(size=(ex.N, 7))
    x[:, 1] *= 1e-30
    x[:, 2] *= 1e30
    x[:, 3] = 0.0
    x[::3, 4] = 0.0
    y = ctx.matvec(x, precision=1)
    ref = H @ x
    for j in range(7):
        if j == 3:
            assert np.all(y[:, j] == 0)
        else:
            assert rel_err(y[:, j], ref[:, j]) < 1e-5, j
    ctx.close()


@pytest.mark.parametrize("name,nstates", [("hbn20", 4), ("random18", 8)])
def test_block_lower_fp32_build(name, nstates, request):
    # without a copy-out the FP32 build keeps only the block-lower part that the split product
    # reads; a solve on the FP32 matrix cores and a later copy-out complete the rest from it
    ex, H = get_case(name, request)
    wr = np.linalg.eigvalsh(H)[:nstates]
    ctx = gpu_context(ex)
    ctx.build(1, scissor=ex.scissor)
    rng = np.random.default_rng(23)
    x = rng.normal(size=(ex.N, 5)) + 1j * rng.normal(size=(ex.N, 5))
    assert rel_err(ctx.matvec(x, precision=1), H @ x) < 1e-5
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=1, tol=1e-6, fp32_arith=0)
    assert np.abs(w - wr).max() < 1e-4
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=1, tol=1e-6, fp32_arith=1)
    assert np.abs(w - wr).max() < 1e-4
    Hg = ctx.get_bse(1).astype(np.complex128)
    assert np.abs(Hg - H).max() < 2e-6 * max(1.0, np.abs(H).max())
    assert np.array_equal(Hg, Hg.conj().T)
    ctx.close()


def test_split_product_blocks_and_chunks():
    # N = 5184: 21 row blocks of 256 (the last one partial) and 3 column chunks of 2048 (the
    # last one ragged), so every kind of work item of the split product (diagonal block,
    # diagonal chunk, full and ragged chunks, transposed blocks) against the FP64 product
    ex = oracle_case(HBN, 72, 1, 0, (1, 1, 10), exchange=True)
    ctx = gpu_context(ex)
    ctx.build(0)
    ctx.build(1)
    rng = np.random.default_rng(17)
    x = rng.normal(size=(ex.N, 37)) + 1j * rng.normal(size=(ex.N, 37))
    y0 = ctx.matvec(x, precision=0)
    y1 = ctx.matvec(x, precision=1)
    for j in range(37):
        assert rel_err(y1[:, j], y0[:, j]) < 1e-5, j
    ctx.close()


@pytest.mark.parametrize("mesh", [24, 45, 72, 96, 120, 128])
def test_convolution_column_lengths(mesh):
    # meshes whose side splits into the larger in-register lengths of the fused column
    # transform (24 = 6 x 4, 45 = 9 x 5, 72 = 9 x 8, 96 = 12 x 8, 120 = 12 x 10,
    # 128 = 16 x 8): the convolution product against the resident FP64 matrix
    print(f"[convolution] hBN mesh {mesh}", flush=True)
    ex = oracle_case(HBN, mesh, 1, 0, (1, 1, 10), exchange=True)
    ctx = gpu_context(ex)
    ctx.build(0)
    rng = np.random.default_rng(mesh)
    x = rng.normal(size=(ex.N, 5)) + 1j * rng.normal(size=(ex.N, 5))
    y = ctx.matvec(x, precision=0)
    assert rel_err(ctx.matvec(x, precision=0, matrix_free=1), y) < 1e-12
    assert rel_err(ctx.matvec(x, precision=1, matrix_free=1), y) < 1e-5
    ctx.close()


MF_FORMS = [1, 2]  # XATU_MATRIX_FREE_CONV (k-space convolution), XATU_MATRIX_FREE_TILES (recomputed tiles)


@pytest.mark.parametrize("form", MF_FORMS)
@pytest.mark.parametrize("name", ["hbn20", "mos2_q_x", "mos2_reduced", "mos2_odd", "random18", "random36"])
def test_matrix_free_products(name, form, request):
    print(f"[matrix-free products] {name} form {form}", flush=True)
    ex, H = get_case(name, request)
    ctx = gpu_context(ex)
    ctx.build(2, scissor=ex.scissor)  # prepare only: nothing resident
    rng = np.random.default_rng(11)
    x = rng.normal(size=(ex.N, 37)) + 1j * rng.normal(size=(ex.N, 37))  # two vector blocks
    y = H @ x
    assert rel_err(ctx.matvec(x, precision=0, matrix_free=form), y) < 1e-12
    assert rel_err(ctx.matvec(x, precision=1, matrix_free=form), y) < 1e-5
    ctx.close()


@pytest.mark.parametrize("form", MF_FORMS)
@pytest.mark.parametrize("name,nstates", [("hbn30_cut20_x", 8), ("random18", 16)])
def test_matrix_free_davidson(name, nstates, form, request):
    print(f"[matrix-free davidson] {name} form {form}", flush=True)
    ex, H = get_case(name, request)
    wr = np.linalg.eigvalsh(H)[:nstates]
    ctx = gpu_context(ex)
    ctx.build(2, scissor=ex.scissor)
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=0, matrix_free=form, tol=1e-7)
    assert rep.converged == 1 and np.abs(w - wr).max() < 1e-6
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=1, matrix_free=form, tol=1e-6)
    assert rep.converged == 1 and np.abs(w - wr).max() < 1e-6
    ctx.close()


# Mesh sides for the whole-plane convolution kernel: 100 = 10 x 10 is the largest FP64 plane
# that fits in LDS (163648 of 163840 bytes); 45 = 9 x 5 an uneven factorisation.
@pytest.mark.parametrize("n", [100, 45])
def test_plane_convolution_sizes(n):
    print(f"[plane convolution] hBN {n}x{n}", flush=True)
    ex = oracle_case(HBN, n, 1, 0, (1, 1, 10), cutoff=20, exchange=True)
    ctx = gpu_context(ex)
    ctx.build(2)  # matrix-free: exchange factors and the diagonal only
    rng = np.random.default_rng(n)
    cols = np.sort(rng.choice(ex.N, size=16, replace=False))
    E = np.zeros((ex.N, 16), complex, order="F")
    E[cols, np.arange(16)] = 1.0
    ref = ex.bse_columns(cols)
    assert np.abs(ctx.matvec(E, precision=0, matrix_free=1) - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.abs(ctx.matvec(E, precision=1, matrix_free=1) - ref).max() <= 1e-5 * np.abs(ref).max()
    ctx.close()


@pytest.mark.parametrize("name,nstates", [("hbn30_cut20_x", 8), ("random18", 16)])
def test_mixed_refines_matrix_free(name, nstates, request):
    # FP32 matrix resident, no FP64 matrix: the FP64 refinement runs matrix-free
    ex, H = get_case(name, request)
    wr = np.linalg.eigvalsh(H)[:nstates]
    ctx = gpu_context(ex)
    ctx.build(1, scissor=ex.scissor)
    w, v, rep = ctx.solve(method=1, nstates=nstates, precision=1, tol=1e-6)
    print(f"[{name}] mixed (matrix-free refinement): err meV {np.round((w - wr) * 1e3, 7)}, "
          f"iters {rep.iterations}, refine {rep.refine_iterations}")
    assert rep.converged == 1 and np.abs(w - wr).max() < 1e-6
    assert np.abs(H @ v - v * w).max() < 1e-5
    ctx.close()
