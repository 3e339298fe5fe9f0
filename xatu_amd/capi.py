"""ctypes binding of the C-ABI in include/xatu_bse.h.

This is the binding a Python caller (or a test) uses to drive the GPU path
directly with its own band stacks; numpy arrays are passed as plain pointers.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libxatu_bse.so"

XATU_OK = 0
XATU_ERR_INVALID_ARGUMENT = 1
XATU_ERR_LOGIC = 2
XATU_ERR_DEVICE = 3
XATU_ERR_OUT_OF_MEMORY = 4
XATU_ERR_NOT_CONVERGED = 5
XATU_ERR_NOT_HERMITIAN = 6
XATU_FP64, XATU_FP32 = 0, 1
XATU_MATRIX_FREE = 2
XATU_MATRIX_FREE_CONV = 1    # matrix-free product forms (solve / matvec)
XATU_MATRIX_FREE_TILES = 2
XATU_POTENTIAL_KELDYSH, XATU_POTENTIAL_COULOMB = 0, 1
XATU_MODE_REALSPACE, XATU_MODE_RECIPROCAL = 0, 1
XATU_ARITH_SPLIT16, XATU_ARITH_FP32 = 0, 1
XATU_SOLVE_DIAG, XATU_SOLVE_DAVIDSON = 0, 1
XATU_PARAM_R0, XATU_PARAM_EPS_M, XATU_PARAM_EPS_S = 0, 1, 2

# every symbol declared in include/xatu_bse.h
EXPORTED = [
    "xatu_create", "xatu_destroy", "xatu_last_error", "xatu_set_lattice", "xatu_set_bands",
    "xatu_fourier_sums", "xatu_build_bse", "xatu_build_bse_basis", "xatu_get_bse", "xatu_free_bse",
    "xatu_solve", "xatu_bse_matvec", "xatu_hf_gradients", "xatu_get_fourier", "xatu_get_cells",
    "xatu_compare_builds", "xatu_get_timings", "xatu_band_gradients",
    "xatu_oscillator_strengths", "xatu_bse_matvec_noalloc",
]


class Interaction(ctypes.Structure):
    _fields_ = [("eps_m", ctypes.c_double), ("eps_s", ctypes.c_double), ("r0", ctypes.c_double),
                ("regularization", ctypes.c_double), ("cutoff", ctypes.c_double), ("ncell", ctypes.c_int),
                ("total_cells", ctypes.c_int), ("potential", ctypes.c_int), ("exchange_potential", ctypes.c_int),
                ("exchange", ctypes.c_int), ("Q", ctypes.c_double * 3), ("mode", ctypes.c_int),
                ("n_reciprocal", ctypes.c_int)]


class SolveOptions(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("precision", ctypes.c_int), ("matrix_free", ctypes.c_int),
                ("nstates", ctypes.c_int), ("block", ctypes.c_int), ("max_basis", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("tol", ctypes.c_double), ("fp32_arith", ctypes.c_int)]


class SolveReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("products", ctypes.c_int), ("restarts", ctypes.c_int),
                ("refine_iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("max_residual", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("ms_products", ctypes.c_double), ("ms_rayleigh_ritz", ctypes.c_double),
                ("ms_projections", ctypes.c_double)]


class Timings(ctypes.Structure):
    _fields_ = [("ms_fourier", ctypes.c_double), ("ms_build", ctypes.c_double),
                ("ms_build_kernel", ctypes.c_double), ("ms_matvec_kernel", ctypes.c_double),
                ("n_matvec_kernel", ctypes.c_int), ("n_matvec_vectors", ctypes.c_int)]


_lib = None
_vp = ctypes.c_void_p


def load(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libxatu_bse.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise ImportError(f"{p} is missing: build it with python -m xatu_amd.build")
        L = ctypes.CDLL(str(p))
        L.xatu_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int]
        L.xatu_destroy.argtypes = [_vp]
        L.xatu_destroy.restype = None
        L.xatu_last_error.argtypes = [_vp]
        L.xatu_last_error.restype = ctypes.c_char_p
        L.xatu_set_lattice.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int]
        L.xatu_set_bands.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
        L.xatu_fourier_sums.argtypes = [_vp, ctypes.POINTER(Interaction)]
        L.xatu_build_bse.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _vp]
        L.xatu_build_bse_basis.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_double, _vp]
        L.xatu_get_bse.argtypes = [_vp, ctypes.c_int, _vp]
        L.xatu_free_bse.argtypes = [_vp, ctypes.c_int]
        L.xatu_solve.argtypes = [_vp, ctypes.POINTER(SolveOptions), _vp, _vp, ctypes.POINTER(SolveReport)]
        L.xatu_bse_matvec.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]
        L.xatu_bse_matvec_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp]
        L.xatu_hf_gradients.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]
        L.xatu_band_gradients.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
        L.xatu_oscillator_strengths.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
        L.xatu_get_fourier.argtypes = [_vp, _vp, _vp]
        L.xatu_get_cells.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), _vp]
        L.xatu_compare_builds.argtypes = [_vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_double)]
        L.xatu_get_timings.argtypes = [_vp, ctypes.POINTER(Timings)]
        _lib = L
    return _lib


class XatuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_vp)


class Context:
    """Owns one xatu_ctx (one device, one HIP stream)."""

    def __init__(self, device: int = 0):
        self.L = load()
        self._h = _vp()
        rc = self.L.xatu_create(ctypes.byref(self._h), device)
        if rc != XATU_OK:
            raise XatuError(rc, "xatu_create failed")
        self.N = 0
        self.natoms = 0
        self.nmesh = 0
        self.ndim = 0

    def close(self):
        if self._h:
            self.L.xatu_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != XATU_OK:
            raise XatuError(rc, self.L.xatu_last_error(self._h).decode())

    def set_lattice(self, bravais, motif_xyz, orb_per_atom, mesh_n: int):
        b = np.ascontiguousarray(bravais, dtype=np.float64)
        m = np.ascontiguousarray(motif_xyz, dtype=np.float64)
        o = np.ascontiguousarray(orb_per_atom, dtype=np.int32)
        self.ndim = b.shape[0]
        self.natoms = m.shape[0]
        self.norb = int(o.sum())
        self.nmesh = mesh_n ** self.ndim
        self._check(self.L.xatu_set_lattice(self._h, self.ndim, _ptr(b), self.natoms, _ptr(m), _ptr(o), mesh_n))

    def set_bands(self, nv, nc, eigval_k, eigvec_k, k_offsets, eigval_kq=None, eigvec_kq=None):
        ek = np.asfortranarray(eigval_k, dtype=np.float64)
        vk = np.asfortranarray(eigvec_k, dtype=np.complex128)
        ekq = None if eigval_kq is None else np.asfortranarray(eigval_kq, dtype=np.float64)
        vkq = None if eigvec_kq is None else np.asfortranarray(eigvec_kq, dtype=np.complex128)
        off = np.ascontiguousarray(k_offsets, dtype=np.int32)
        nk = ek.shape[1]
        self.N = nk * nv * nc
        self.nv, self.nc = nv, nc
        self.nk, self.nb = nk, nv + nc
        self._keep = (ek, vk, ekq, vkq, off)
        self._check(self.L.xatu_set_bands(self._h, nk, nv, nc, _ptr(ek), _ptr(ekq), _ptr(vk), _ptr(vkq),
                                          _ptr(off)))

    def fourier_sums(self, eps_m, eps_s, r0, cutoff, ncell, total_cells, regularization=0.0,
                     potential=XATU_POTENTIAL_KELDYSH, exchange_potential=XATU_POTENTIAL_KELDYSH,
                     exchange=False, Q=(0.0, 0.0, 0.0), mode=0, n_reciprocal=1):
        it = Interaction(eps_m, eps_s, r0, regularization, cutoff, ncell, total_cells, potential,
                         exchange_potential, int(bool(exchange)), (ctypes.c_double * 3)(*Q), int(mode),
                         int(n_reciprocal))
        self._check(self.L.xatu_fourier_sums(self._h, ctypes.byref(it)))

    def build(self, precision=XATU_FP64, scissor=0.0, copy_out=False):
        out = None
        if copy_out:
            dt = np.complex128 if precision == XATU_FP64 else np.complex64
            out = np.zeros((self.N, self.N), dtype=dt, order="F")
        self._check(self.L.xatu_build_bse(self._h, precision, scissor, _ptr(out)))
        return out

    def build_basis(self, basis, scissor=0.0):
        b = np.ascontiguousarray(basis, dtype=np.int32)
        n = b.shape[0]
        out = np.zeros((n, n), dtype=np.complex128, order="F")
        self._check(self.L.xatu_build_bse_basis(self._h, n, _ptr(b), scissor, _ptr(out)))
        return out

    def get_bse(self, precision=XATU_FP64):
        dt = np.complex128 if precision == XATU_FP64 else np.complex64
        out = np.zeros((self.N, self.N), dtype=dt, order="F")
        self._check(self.L.xatu_get_bse(self._h, precision, _ptr(out)))
        return out

    def free_bse(self, precision):
        self._check(self.L.xatu_free_bse(self._h, precision))

    def solve(self, method=XATU_SOLVE_DAVIDSON, nstates=8, precision=XATU_FP64, matrix_free=False, block=0,
              max_basis=0, max_iter=0, tol=0.0, want_vectors=True, fp32_arith=0):
        opt = SolveOptions(method, precision, int(matrix_free), nstates, block, max_basis, max_iter, tol,
                           int(fp32_arith))
        n_out = self.N if method == XATU_SOLVE_DIAG and nstates <= 0 else nstates
        w = np.zeros(max(n_out, 1))
        v = np.zeros((self.N, n_out), dtype=np.complex128, order="F") if want_vectors else None
        rep = SolveReport()
        self._check(self.L.xatu_solve(self._h, ctypes.byref(opt), _ptr(w), _ptr(v), ctypes.byref(rep)))
        return w[:n_out], v, rep

    def matvec(self, x, precision=XATU_FP64, matrix_free=False):
        x = np.asfortranarray(x, dtype=np.complex128)
        if x.ndim == 1:
            x = x[:, None]
        y = np.zeros_like(x, order="F")
        self._check(self.L.xatu_bse_matvec(self._h, precision, int(matrix_free), x.shape[1], _ptr(x),
                                           _ptr(y)))
        return y

    def matvec_noalloc(self, x_dev: int, y_dev: int, nvec: int):
        """y = H x with the resident FP64 matrix, x and y device addresses of N x nvec complex128
        column-major buffers (no allocation inside the call)."""
        self._check(self.L.xatu_bse_matvec_noalloc(self._h, int(nvec), _vp(x_dev), _vp(y_dev)))

    def hf_gradients(self, nstates, params, eigvec=None):
        p = np.ascontiguousarray(params, dtype=np.int32)
        v = None if eigvec is None else np.asfortranarray(eigvec, dtype=np.complex128)
        g = np.zeros((nstates, len(p)), order="F")
        self._check(self.L.xatu_hf_gradients(self._h, nstates, _ptr(v), len(p), _ptr(p), _ptr(g)))
        return g

    def band_gradients(self, nstates, eigvec=None):
        """(w, g): w[n, stack, k, b] = dE_n/d eps_b(k); g[n, stack, k, b, orb] = dE_n/d conj C_b(k)[orb]
        (stack 0: k stack, 1: k+Q stack)."""
        v = None if eigvec is None else np.asfortranarray(eigvec, dtype=np.complex128)
        w = np.zeros((nstates, 2, self.nk, self.nb))
        g = np.zeros((nstates, 2, self.nk, self.nb, self.norb), dtype=np.complex128)
        self._check(self.L.xatu_band_gradients(self._h, nstates, _ptr(v), _ptr(w), _ptr(g)))
        return w, g

    def oscillator_strengths(self, velocity, nstates, eigvec=None):
        """out[j, n] = sum_i X_n(i) velocity[i, j] (velocity: N x 3 complex)."""
        vel = np.asfortranarray(velocity, dtype=np.complex128)
        v = None if eigvec is None else np.asfortranarray(eigvec, dtype=np.complex128)
        out = np.zeros((3, nstates), dtype=np.complex128, order="F")
        self._check(self.L.xatu_oscillator_strengths(self._h, nstates, _ptr(v), _ptr(vel), _ptr(out)))
        return out

    def get_fourier(self):
        ft = np.zeros((self.natoms, self.natoms, self.nmesh), dtype=np.complex128, order="F")
        ftq = np.zeros((self.natoms, self.natoms), dtype=np.complex128, order="F")
        self._check(self.L.xatu_get_fourier(self._h, _ptr(ft), _ptr(ftq)))
        return ft, ftq

    def get_cells(self):
        n = ctypes.c_int(0)
        self._check(self.L.xatu_get_cells(self._h, ctypes.byref(n), None))
        c = np.zeros((n.value, self.ndim), dtype=np.int32)
        self._check(self.L.xatu_get_cells(self._h, ctypes.byref(n), _ptr(c)))
        return c

    def compare_builds(self, kblock):
        """(per-block max |H32 - H64| as an nblk x nblk array, overall max)."""
        rows = kblock * self.nv * self.nc
        nblk = (self.N + rows - 1) // rows
        blk = np.zeros((nblk, nblk), order="F")
        mx = ctypes.c_double(0.0)
        self._check(self.L.xatu_compare_builds(self._h, kblock, _ptr(blk), ctypes.byref(mx)))
        return blk, mx.value

    def timings(self) -> Timings:
        t = Timings()
        self._check(self.L.xatu_get_timings(self._h, ctypes.byref(t)))
        return t
