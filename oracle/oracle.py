"""CPU FP64 oracle for the Xatu BSE hot path -- TEST INFRASTRUCTURE ONLY.

Used by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` as the checker of the HIP path; the product never imports it.

Pieces and the reference code they restate (file:line in alejandrojuria/xatu):

- ``parse_model``          src/ConfigurationBase.cpp:43-147, src/SystemConfiguration.cpp:33-228
- ``System``               src/Lattice.cpp:56-131, 140-279 (reciprocal lattice, MP / reduced /
                           shifted meshes, findEquivalentPointBZ), src/SystemTB.cpp:60-125
                           (H(k) and eig_sym -> numpy.linalg.eigh = LAPACK zheevd)
- ``fix_global_phase``     src/ExcitonTB.cpp:673-683
- ``Exciton``              src/ExcitonTB.cpp:20-34, 128-144 (bands), 707-815
                           (initializeResultsH0, initializeHamiltonian), 840-921 (BShamiltonian,
                           through liboracle.so), 930-959 (diag = numpy eigh)
- ``davidson_reference``   src/davidson.cpp:5-61, restated line by line (including its
                           identity guess and dense (M - theta I) corrections)
- ``liboracle.so``         oracle.cpp: potentials, lattice sums and the per-element BSE loop.

Parity status: every function above is pinned by the reference's regression values
(test/tests.cpp) and by closed forms / invariants in tests/test_oracle.py, except
the explicit dV/dr0 kind of ``ft_stack`` (no reference counterpart), which is pinned
by central finite differences of the potential in the same tests.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import re

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

POT_KELDYSH, POT_COULOMB, POT_KELDYSH_DR0, POT_KELDYSH_DEPS = 0, 1, 2, 3
_POT_NAMES = {"keldysh": POT_KELDYSH, "coulomb": POT_COULOMB}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    """Load (building if needed) oracle/liboracle.so."""
    global _LIB
    if _LIB is None:
        so = _HERE / "liboracle.so"
        if not so.exists():
            import subprocess
            subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
        L = ctypes.CDLL(str(so))
        L.oracle_stvh0.restype = ctypes.c_double
        L.oracle_stvh0.argtypes = [ctypes.c_double]
        L.oracle_y0.restype = ctypes.c_double
        L.oracle_y0.argtypes = [ctypes.c_double]
        L.oracle_h1.restype = ctypes.c_double
        L.oracle_h1.argtypes = [ctypes.c_double]
        L.oracle_keldysh_dxf.restype = ctypes.c_double
        L.oracle_keldysh_dxf.argtypes = [ctypes.c_double]
        L.oracle_potential.restype = ctypes.c_double
        L.oracle_potential.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6
        L.oracle_truncate_supercell.restype = ctypes.c_int
        L.oracle_truncate_supercell.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int]
        L.oracle_motif_ft.restype = None
        L.oracle_motif_ft.argtypes = ([ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int]
                                      + [ctypes.c_double] * 6 + [_dp])
        L.oracle_ft_stack.restype = None
        L.oracle_ft_stack.argtypes = ([ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int]
                                      + [ctypes.c_double] * 6 + [_dp])
        L.oracle_find_equivalent.restype = ctypes.c_int
        L.oracle_find_equivalent.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        bse_args = [ctypes.c_int, _ip, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                    _dp, ctypes.c_int, _ip, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp,
                    _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_bse.restype = None
        L.oracle_bse.argtypes = bse_args + [_dp]
        L.oracle_bse_columns.restype = None
        L.oracle_bse_columns.argtypes = bse_args + [ctypes.c_int, _ip, _dp]
        L.oracle_bse_apply.restype = None
        L.oracle_bse_apply.argtypes = bse_args + [ctypes.c_int, _dp, _dp]
        _LIB = L
    return _LIB


def _d(a):
    return np.ascontiguousarray(a).ctypes.data_as(_dp)


# --------------------------------------------------------------------------- parsing

def _blocks(path):
    """'#' argument blocks of a Xatu text file (ConfigurationBase.cpp:57-119)."""
    text = Path(path).read_text()
    ends_nl = text.endswith("\n")
    lines = text.split("\n")
    if ends_nl:
        lines = lines[:-1]
    found, contents = [], {}
    for i, line in enumerate(lines):
        if "#" in line and i + 1 < len(lines):
            found.append("".join(line[line.index("#") + 1:].split()).lower())
    arg, content = "", []
    for i, line in enumerate(lines):
        if "#" in line:
            if content:
                contents[arg] = content
                content = []
            arg = "".join(line[line.index("#") + 1:].split()).lower()
        elif not line or "!" in line:
            pass
        else:
            content.append(line.replace(",", " ").replace(";", " "))
        if i + 1 == len(lines) and not ends_nl:
            contents[arg] = content
    return found, contents


def _matrices(content):
    cplx = ("i" in content[0]) or ("j" in content[0])
    dim = 0
    for line in content:
        if "&" in line:
            break
        dim += 1
    mats, cur, row = [], np.zeros((dim, dim), complex), 0
    for line in content:
        if "&" in line:
            mats.append(cur)
            cur, row = np.zeros((dim, dim), complex), 0
            continue
        toks = line.split()
        vals = []
        if cplx:
            for t in range(0, len(toks), 2):
                vals.append(float(toks[t]) + 1j * float(toks[t + 1].rstrip("ij")))
        else:
            vals = [float(t) for t in toks]
        cur[row, :len(vals)] = vals
        row += 1
    return np.array(mats)


def parse_model(path):
    found, c = _blocks(path)
    m = {}
    for arg in found:
        content = c.get(arg, [])
        if arg == "dimension":
            m["ndim"] = int(content[0].split()[0])
        elif arg == "bravaislattice":
            m["bravais"] = np.array([[float(x) for x in l.split()] for l in content])
        elif arg == "motif":
            m["motif"] = np.array([[float(x) for x in l.split()[:4]] for l in content])
        elif arg == "norbitals":
            m["norbitals"] = [int(x) for x in content[0].split()]
        elif arg == "filling":
            m["filling"] = int(float(content[0].split()[0]))
        elif arg == "bravaisvectors":
            m["cells"] = np.array([[float(x) for x in l.split()] for l in content])
        elif arg == "hamiltonian":
            m["H"] = _matrices(content)
        elif arg == "overlap":
            m["S"] = _matrices(content)
    return m


# --------------------------------------------------------------------------- system

def fix_global_phase(V):
    """Columns scaled so that the sum of each column is real and >= 0."""
    s = V.sum(axis=-2, keepdims=True)
    return V * np.exp(-1j * np.angle(s))


class System:
    def __init__(self, model):
        if isinstance(model, (str, Path)):
            model = parse_model(model)
        self.ndim = model["ndim"]
        self.bravais = np.asarray(model["bravais"], float)
        self.motif = np.asarray(model["motif"], float)
        self.natoms = len(self.motif)
        self.orbitals = model["norbitals"]
        self.cells = np.asarray(model["cells"], float)
        self.H = np.asarray(model["H"], complex)
        self.S = model.get("S")
        self.filling = model["filling"]
        # SystemTB.cpp:30 / utils.cpp:176-178: H(0) holding only one triangle marks the
        # matrices as triangles (H(k) = T(k) + T(k)^H - diag T(k))
        H0 = self.H[0]
        self.triangular = bool(np.all(np.tril(H0, -1) == 0) != np.all(np.triu(H0, 1) == 0))
        self.au = False
        self.fermi = self.filling - 1
        self.orb_per_atom = [self.orbitals[int(s)] for s in self.motif[:, 3]]
        self.basisdim = sum(self.orb_per_atom)
        self.a = float(np.linalg.norm(self.bravais[0]))
        if self.ndim == 1:
            self.recip = 2 * np.pi * self.bravais / np.linalg.norm(self.bravais) ** 2
        else:
            A = self.bravais[:, :self.ndim]
            self.recip = np.zeros((self.ndim, 3))
            for i in range(self.ndim):
                e = np.zeros(self.ndim)
                e[i] = 2 * np.pi
                self.recip[i, :self.ndim] = np.linalg.solve(A, e)
        self.inv = np.linalg.inv(self.recip @ self.recip.T)
        self.factor = 1
        self.kpoints = None
        self.meshBZ = None

    # Lattice.cpp:293-312 (first coefficient fastest)
    def combinations(self, n, centered=False):
        shift = n // 2 if centered else 0
        grids = np.meshgrid(*[np.arange(n) - shift] * self.ndim, indexing="ij")
        return np.stack([g.ravel(order="F") for g in grids], axis=1).astype(float)

    def mp_mesh(self, n, factor=1):
        c = self.combinations(n)
        if n % 2 == 1:
            c = c + 0.5
        return ((2 * c - n) / (2 * n * factor)) @ self.recip

    def brillouin_zone_mesh(self, n):
        self.kpoints = self.mp_mesh(n)
        self.meshBZ = self.kpoints.copy()

    def reduced_brillouin_zone_mesh(self, n, factor):
        self.brillouin_zone_mesh(n * factor)
        self.kpoints = self.mp_mesh(n, factor)
        self.factor = factor

    def shift_bz(self, shift):
        self.kpoints = self.kpoints + np.asarray(shift, float)

    def truncate_supercell(self, ncell, radius):
        out = np.zeros((ncell ** self.ndim, 3))
        n = lib().oracle_truncate_supercell(self.ndim, _d(self.bravais), ncell, radius, _d(out), len(out))
        return out[:n].copy()

    def _bloch(self, k, M):
        k = np.atleast_2d(k)
        ph = np.exp(1j * (k @ self.cells.T))  # nk x ncells
        h = np.einsum("kr,rij->kij", ph, M)
        if self.triangular:  # SystemTB.cpp:69-72
            d = np.einsum("kii->ki", h)
            h = h - np.einsum("ki,ij->kij", d / 2, np.eye(h.shape[-1]))
            h = h + np.conj(np.swapaxes(h, -1, -2))
        return h

    def hamiltonian(self, k):
        return self._bloch(k, self.H)

    def solve_bands(self, k):
        h = self.hamiltonian(k)
        if self.S is not None and len(self.S):
            if self.au:  # SystemTB.cpp:117-122 (Hartree -> eV with the reference's 27.2)
                h = h * 27.2
            s = self._bloch(k, self.S)
            w, v = np.linalg.eigh(s)
            root = np.einsum("kij,kj,klj->kil", v, 1 / np.sqrt(w), v.conj())
            h = root @ h @ root
        return np.linalg.eigh(h)

    def find_equivalent(self, k, ncell):
        k = np.asarray(k, float)
        return lib().oracle_find_equivalent(self.ndim, _d(self.recip), _d(self.inv), _d(k), ncell, self.factor)


# --------------------------------------------------------------------------- exciton

def atomic_gauge(system, stack):
    """SystemTB::latticeToAtomicGauge (src/SystemTB.cpp:166-182) on every state of a band
    stack (norb x nb x nk): coefficient alpha at k-point k times exp(-i k . tau_atom(alpha)).
    BShamiltonian applies it with the k-point of the row to the k and k+Q stacks alike
    (src/ExcitonTB.cpp:877-884)."""
    tau = np.repeat(system.motif[:, :3], system.orb_per_atom, axis=0)   # norb x 3
    ph = np.exp(-1j * (system.kpoints @ tau.T))                           # nk x norb
    return np.asfortranarray(stack * ph.T[:, None, :])


class Exciton:
    """The reference ExcitonTB pipeline on the CPU (realspace mode; lattice gauge, or the
    atomic gauge with gauge = "atomic")."""

    def __init__(self, system, ncell, nbands=1, nrmbands=0, parameters=(1, 5, 1), Q=(0, 0, 0), bands=None):
        self.sys = system if isinstance(system, System) else System(system)
        self.ncell = ncell
        self.eps_m, self.eps_s, self.r0 = map(float, parameters)
        if self.r0 == 0:
            raise ValueError("Error: r0 must be non-zero")
        self.Q = np.asarray(Q, float)
        self.cutoff = ncell / 2.5
        self.scissor = 0.0
        self.exchange = False
        self.potential = "keldysh"
        self.exchange_potential = "keldysh"
        self.regularization = 0.0
        self.gauge = "lattice"
        f = self.sys.fermi
        if bands is None:
            self.valence = list(range(f - nbands - nrmbands + 1, f - nrmbands + 1))
            self.conduction = list(range(f + 1 + nrmbands, f + nbands + nrmbands + 1))
        else:
            self.valence = [b + f for b in bands if b <= 0]
            self.conduction = [b + f for b in bands if b > 0]
        self.band_list = self.valence + self.conduction

    def initialize(self):
        s = self.sys
        if s.kpoints is None:
            raise ValueError("Error: BZ mesh must be initialized first")
        if self.regularization < 1e-10:
            self.regularization = s.a
        self.nk = len(s.kpoints)
        self.nv, self.nc = len(self.valence), len(self.conduction)
        self.N = self.nk * self.nv * self.nc
        self.total_cells = (self.ncell * s.factor) ** s.ndim
        # basis rows {vcol, ccol, k}, order k > c > v (Exciton.tpp:128-147)
        b = []
        for k in range(self.nk):
            for ic in range(self.nc):
                for iv in range(self.nv):
                    b.append((iv, self.nv + ic, k))
        self.basis = np.array(b, dtype=np.int32)
        bl = np.array(self.band_list)
        w, v = s.solve_bands(s.kpoints)
        v = fix_global_phase(v)
        self.eigval_k = np.asfortranarray(w[:, bl].T)                    # nb x nk
        self.eigvec_k = np.asfortranarray(np.transpose(v[:, :, bl], (1, 2, 0)))  # norb x nb x nk
        if np.linalg.norm(self.Q) != 0:
            w, v = s.solve_bands(s.kpoints + self.Q)
            v = fix_global_phase(v)
            self.eigval_kq = np.asfortranarray(w[:, bl].T)
            self.eigvec_kq = np.asfortranarray(np.transpose(v[:, :, bl], (1, 2, 0)))
        else:
            self.eigval_kq = self.eigval_k.copy(order="F")
            self.eigvec_kq = self.eigvec_k.copy(order="F")
        self.cells = s.truncate_supercell(self.ncell, s.a * self.cutoff)
        self.ft = self.ft_stack(_POT_NAMES[self.potential])
        self.ftq = self.ft_at(self.Q, _POT_NAMES[self.exchange_potential]) if self.exchange \
            else np.zeros((s.natoms, s.natoms), complex, order="F")

    def _pot_args(self):
        return (self.eps_m, self.eps_s, self.r0, self.regularization, self.sys.a * self.cutoff + 1e-5,
                float(self.total_cells))

    def ft_stack(self, kind):
        s = self.sys
        mesh = np.ascontiguousarray(s.meshBZ)
        out = np.zeros((s.natoms, s.natoms, len(mesh)), complex, order="F")
        lib().oracle_ft_stack(s.natoms, _d(s.motif[:, :3]), len(self.cells), _d(self.cells), len(mesh), _d(mesh),
                              kind, *self._pot_args(), out.ctypes.data_as(_dp))
        return out

    def ft_at(self, k, kind):
        s = self.sys
        out = np.zeros((s.natoms, s.natoms), complex, order="F")
        lib().oracle_motif_ft(s.natoms, _d(s.motif[:, :3]), len(self.cells), _d(self.cells),
                              _d(np.asarray(k, float)), kind, *self._pot_args(), out.ctypes.data_as(_dp))
        return out

    def _bse_args(self, ft=None, ftq=None, eigval_k=None, eigval_kq=None, scissor=None, eigvec_k=None,
                  eigvec_kq=None, exchange=None):
        """Arguments of the C restatement of the element rule, plus the arrays they point into."""
        s = self.sys
        ft = self.ft if ft is None else ft
        ftq = self.ftq if ftq is None else ftq
        ek = self.eigval_k if eigval_k is None else eigval_k
        ekq = self.eigval_kq if eigval_kq is None else eigval_kq
        vk = self.eigvec_k if eigvec_k is None else eigvec_k
        vkq = self.eigvec_kq if eigvec_kq is None else eigvec_kq
        if self.gauge == "atomic":
            vk, vkq = atomic_gauge(s, vk), atomic_gauge(s, vkq)
        sc = self.scissor if scissor is None else scissor
        ex = self.exchange if exchange is None else exchange
        keep = [np.repeat(np.arange(s.natoms), s.orb_per_atom).astype(np.int32), np.ascontiguousarray(s.kpoints)]
        keep += [np.asfortranarray(a) for a in (vk, vkq, ek, ekq, ft, ftq)]
        orb_atom, kp, vk_, vkq_, ek_, ekq_, ft_, ftq_ = keep
        ptr = lambda a: a.ctypes.data_as(_dp)  # noqa: E731
        args = (self.N, self.basis.ctypes.data_as(_ip), self.nk, ptr(kp), s.basisdim, len(self.band_list), ptr(vk_),
                ptr(vkq_), ptr(ek_), ptr(ekq_), s.natoms, orb_atom.ctypes.data_as(_ip), ptr(ft_), ptr(ftq_),
                int(bool(ex)), float(sc), s.ndim, _d(s.recip), _d(s.inv), self.ncell, s.factor)
        return args, keep

    def bse(self, **kw):
        """Full HBS (+ conjugate lower triangle) from the restated per-element loop. Keywords
        replace the stacks, lattice sums, scissor or exchange flag (see _bse_args)."""
        args, keep = self._bse_args(**kw)
        out = np.zeros((self.N, self.N), complex, order="F")
        lib().oracle_bse(*args, out.ctypes.data_as(_dp))
        return out

    def bse_columns(self, cols, **kw):
        """Columns cols of HBS (N x len(cols)) from the same element rule."""
        args, keep = self._bse_args(**kw)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.zeros((self.N, len(c)), complex, order="F")
        lib().oracle_bse_columns(*args, len(c), c.ctypes.data_as(_ip), out.ctypes.data_as(_dp))
        return out

    def bse_apply(self, x, **kw):
        """HBS x for x of shape N x nvec, every element evaluated by the element rule (no matrix
        is stored)."""
        args, keep = self._bse_args(**kw)
        x = np.asfortranarray(x, dtype=complex)
        if x.ndim == 1:
            x = x[:, None]
        y = np.zeros_like(x, order="F")
        lib().oracle_bse_apply(*args, x.shape[1], x.ctypes.data_as(_dp), y.ctypes.data_as(_dp))
        return y

def davidson_reference(M, neigval, tol=1e-8):
    """davidson_method restated (src/davidson.cpp:5-61); returns (eigval[:neigval], eigvec (N x 2 neigval))."""
    n = M.shape[0]
    if not np.allclose(M, M.conj().T, atol=0, rtol=0):
        raise ValueError("davidson_method: provided matrix must be hermitian")
    max_iterations = n // 2
    k = 2 * neigval
    guess = np.eye(n, k, dtype=complex)
    aux = np.ones(neigval)
    eigval = eigvec = None
    for i in range(max_iterations):
        Q, _ = np.linalg.qr(guess, mode="reduced")
        P = Q.conj().T @ M @ Q
        eigval, v = np.linalg.eigh(P)
        eigvec = Q[:, :(i + 1) * k] @ v[:, :k]
        new = []
        for j in range(k):
            w = (M - eigval[j] * np.eye(n)) @ eigvec[:, j]
            new.append(w / (eigval[j] - M[j, j]))
        Q = np.concatenate([Q, np.array(new).T], axis=1)
        if np.linalg.norm(eigval[:neigval] - aux) < tol:
            break
        aux = eigval[:neigval].copy()
        guess = Q
    return eigval[:neigval], eigvec


def detect_degeneracies(eigval, n, precision=6):
    """src/utils.cpp:121-152."""
    pairs, prev, deg = [], eigval[0], 1
    thr = 10.0 ** (-precision)
    for i in range(1, n):
        e = eigval[i]
        if abs(e - prev) < thr:
            deg += 1
        else:
            pairs.append((prev, deg))
            deg = 1
        prev = e
    pairs.append((prev, deg))
    return pairs


def array2hash(a, aliased=False):
    """Test fingerprint (include/xatu/utils.hpp:61-78); rows/cols as Armadillo sees them.

    ``aliased=True`` reproduces a call on an Armadillo subview (e.g. ``mat.row(i)``):
    there ``T copyArray(array)`` is another view of the same memory, so the scaled
    values also feed the |.| and non-zero terms.
    """
    a = np.asarray(a)
    if a.ndim == 1:
        a = a[:, None]
    i = np.arange(a.shape[0])[:, None]
    j = np.arange(a.shape[1])[None, :]
    c = a * (i * j + j + 1)
    base = c if aliased else a
    return float((np.abs(base).sum() + np.count_nonzero(base) + (np.real(c) * 1.5).sum()
                  + (np.imag(c) * 1.5).sum() + (np.imag(c) * np.real(c)).sum()) / a.size)


# --------------------------------------------------------------------------- optics

BOHR = 0.52917721067121   # Angstrom (src/skubo_w.f90:229-237)
HARTREE = 27.211385       # eV


def velocity_matrix_elements(ex):
    """Single-particle velocity matrix elements of the k-stack bands, in atomic units,
    as exciton_oscillator_strength computes them (src/skubo_w.f90:175-330 with
    hoppings_observables_TB :407-441, get_vme_kernels :445-527, get_eigen_vme :531-627).

    Returns V[k, j, n, m] (j = x, y, z; n, m band columns of the stacks). Kept as the
    reference has it: only the lower orbital triangle of the kernels is summed and the
    upper one mirrored (so the orbital diagonal of the derivative kernel is
    conjugated), the position term uses the motif only at the first listed cell, the
    overlap is the identity at every cell, z components of R are dropped, and the band
    diagonal of V is conjugated by the same mirroring.
    """
    s = ex.sys
    if s.S is not None and len(s.S):
        raise NotImplementedError("velocity matrix elements with overlap matrices")
    norb = s.basisdim
    Rv = s.cells.copy() / BOHR
    Rv[:, 2] = 0.0
    B = np.repeat(s.motif[:, :3], s.orb_per_atom, axis=0) / BOHR
    hh = s.H / HARTREE                                   # [R, a, a']
    kp = s.kpoints * BOHR
    lower = np.tril(np.ones((norb, norb), bool))
    C = ex.eigvec_k                                      # norb x nb x nk
    E = ex.eigval_k / HARTREE                            # nb x nk
    nb = C.shape[1]
    V = np.zeros((ex.nk, 3, nb, nb), complex)
    for k in range(ex.nk):
        f = np.exp(1j * (kp[k, 0] * Rv[:, 0] + kp[k, 1] * Rv[:, 1]))   # phase per cell
        hder = np.zeros((3, norb, norb), complex)
        sder = np.zeros((3, norb, norb), complex)
        aker = np.zeros((3, norb, norb), complex)
        for j in range(2):
            hder[j] = np.einsum("r,rab->ab", f * 1j * Rv[:, j], hh)
            sder[j] = np.sum(f * 1j * Rv[:, j]) * np.eye(norb)
        for j in range(3):
            aker[j] = f[0] * np.diag(B[:, j]) + 1j * sder[j]
        hder = np.where(lower, hder, 0)
        sder = np.where(lower, sder, 0)
        aker = np.where(lower, aker, 0)
        # mirror the lower triangle (the diagonal is overwritten by its own image)
        hder_f = np.where(lower, hder, 0) + 0j
        aker_f = aker.copy()
        for j in range(3):
            hl, al, sl = hder[j], aker[j], sder[j]
            hu = hl.conj().T
            au = al.conj().T + 1j * sl.conj().T
            hder_f[j] = np.where(lower.T, hu, hl)
            aker_f[j] = np.where(lower.T, au, al)
        c = C[:, :, k]
        e = E[:, k]
        for j in range(3):
            va = c.conj().T @ hder_f[j] @ c                               # [n, m]
            t1 = c.conj().T @ aker_f[j] @ c                               # sum conj(c_n) c_m A_ab
            t2 = c.conj().T @ aker_f[j].conj().T @ c                      # sum conj(c_n) c_m conj(A_ba)
            vb = 1j * (e[:, None] * t1 - e[None, :] * t2)
            v = va + vb
            low = np.tril(np.ones((nb, nb), bool))                        # computed for m <= n
            V[k, j] = np.where(low.T, v.conj().T, v)
    return V


def exciton_oscillator_strength(ex, eigvec):
    """vme_ex(j, n) = sum_{k,c,v} X_n(k c v) V[k, j, v, nv + c] (src/skubo_w.f90:300-323,
    ResultTB::excitonOscillatorStrength src/ResultTB.cpp:242-298): 3 x nstates, a.u."""
    V = velocity_matrix_elements(ex)
    nv, nc = ex.nv, ex.nc
    pair = np.zeros((ex.nk, nc, nv, 3), complex)
    for ic in range(nc):
        for iv in range(nv):
            pair[:, ic, iv, :] = V[:, :, iv, nv + ic]
    return pair.reshape(ex.N, 3).T @ np.asarray(eigvec)


def exciton_spectrum(vme_ex, e_ex, nk, vcell_A2, w0=0.0, wrange=6.0, nw=2000, eta=0.05):
    """Lorentzian-broadened exciton conductivity (src/skubo_w.f90:95-130, broad_vector
    :684-720): returns (w in eV, sigma[w, j, j'] for j, j' in {x, y}), with
    skubo(j,j',n) = pi/(nk vcell) conj(v_jn) v_j'n / e_n in atomic units."""
    e = np.asarray(e_ex) / HARTREE
    vcell = vcell_A2 / BOHR ** 2
    wp = (w0 + wrange / nw * np.arange(nw)) / HARTREE
    et = eta / HARTREE
    sk = np.pi / (nk * vcell) * np.einsum("jn,kn->jkn", vme_ex.conj(), vme_ex) / e
    L = (1 / np.pi) * np.imag(1.0 / (wp[:, None] - e[None, :] - 1j * et))   # [w, n]
    sig = np.einsum("wn,jkn->wjk", L, np.real(sk[:2, :2]))
    return wp * HARTREE, sig


# --------------------------------------------------------------------------- reciprocal-space mode

EC = 1.6021766E-19    # include/xatu/ExcitonTB.hpp:10-14
EPS0 = 8.8541878E-12


def truncate_reciprocal_supercell(system, ncell, radius):
    """Lattice::truncateReciprocalSupercell (src/Lattice.cpp:349-370): centred combinations
    (first coefficient fastest) of the reciprocal vectors with |G| < radius + 1e-5."""
    combos = system.combinations(ncell, centered=True)
    G = combos @ system.recip
    return G[np.linalg.norm(G, axis=1) < radius + 1e-5]


def unit_cell_area(system):
    """Lattice::computeUnitCellArea (src/Lattice.cpp:119-131)."""
    b = system.bravais
    if system.ndim == 1:
        return float(np.linalg.norm(b[0]))
    c = np.cross(b[0], b[1])
    return float(np.linalg.norm(c)) if system.ndim == 2 else float(np.dot(c, b[2]))


def keldysh_ft(ex, q):
    """ExcitonTB::keldyshFT (src/ExcitonTB.cpp:408-426), vectorised over the rows of q."""
    s = ex.sys
    b1 = float(np.linalg.norm(s.recip[0]))
    qn = np.linalg.norm(np.atleast_2d(q), axis=1)
    eps_bar = (ex.eps_m + ex.eps_s) / 2
    with np.errstate(divide="ignore"):
        pot = np.where(qn < b1 / ex.total_cells, 0.0, 1.0 / (qn * (1 + ex.r0 * qn)))
    return pot * EC * 1E10 / (2 * EPS0 * eps_bar * unit_cell_area(s) * ex.total_cells)


def bse_reciprocal(ex, n_reciprocal):
    """BShamiltonian in reciprocal-space mode (src/ExcitonTB.cpp:562-621, 895-916):
    D = sum_G Ic conj(Iv) W(k - k' + G) with the Bloch coherence factors
    Ic = sum_al conj(C_kc,al) C_k'c',al e^{i(k-k'+G).tau_al} (conduction, k+Q stack) and Iv
    likewise for the valence pair; X = sum_G Ic' conj(Iv') W(Q + G) with
    Ic' = sum_al conj(C_kc,al) V_kv,al e^{i(Q+G).tau_al}, Iv' the same at k'. Upper triangle,
    diagonal rule and conjugate mirror as in realspace."""
    s = ex.sys
    G = truncate_reciprocal_supercell(s, n_reciprocal, ex.cutoff * float(np.linalg.norm(s.recip[0])))
    tau = np.repeat(s.motif[:, :3], s.orb_per_atom, axis=0)            # norb x 3
    kp = s.kpoints
    nv, nc, nk = ex.nv, ex.nc, ex.nk
    N = ex.N
    ki = np.repeat(np.arange(nk), nc * nv)
    ci = np.tile(np.repeat(np.arange(nc), nv), nk)
    vi = np.tile(np.arange(nv), nk * nc)
    C = ex.eigvec_kq[:, nv:, :]                                          # norb x nc x nk
    V = ex.eigvec_k[:, :nv, :]
    Cr = C[:, ci, ki]                                                    # norb x N (conduction of row)
    Vr = V[:, vi, ki]
    H = np.zeros((N, N), complex)
    for i in range(N):
        j = np.arange(i, N)
        q = kp[ki[i]] - kp[ki[j]]                                        # (k - k') per column
        D = np.zeros(len(j), complex)
        X = np.zeros(len(j), complex)
        for g in G:
            ph = np.exp(1j * ((q + g) @ tau.T))                          # ncol x norb
            Ic = ((Cr[:, i].conj()[:, None] * Cr[:, j]).T * ph).sum(1)
            Iv = ((Vr[:, i].conj()[:, None] * Vr[:, j]).T * ph).sum(1)
            D += Ic * np.conj(Iv) * keldysh_ft(ex, q + g)
            if ex.exchange:
                phq = np.exp(1j * ((ex.Q + g) @ tau.T))                 # norb
                Icx = np.sum(Cr[:, i].conj() * Vr[:, i] * phq)
                Ivx = ((Cr[:, j].conj() * Vr[:, j]).T * phq).sum(1)
                X += Icx * np.conj(Ivx) * keldysh_ft(ex, ex.Q + g)[0]
        F = D - X
        e = (ex.scissor + ex.eigval_kq[nv + ci[i], ki[i]]) - ex.eigval_k[vi[i], ki[i]]
        H[i, j] = -F
        h = e / 2 - F[0] / 2
        H[i, i] = h + h
    H = np.triu(H, 1) + np.triu(H, 1).conj().T + np.diag(np.real(np.diag(H)))
    return H


# ------------------------------------------------------------------ result observables

def exciton_spin(ex, x):
    """ResultTB::spinX (src/ResultTB.cpp:21-99): <S_z> of the state x as {total, hole,
    electron}. Orbital o of an atom carries s_z = +1/2 (o even) or -1/2 (o odd); per k the
    hole matrix is Sh(i, j) = sum_a conj(V_j(a)) s_a V_i(a) (k stack, valence) and the
    electron matrix Se(i, j) = sum_a conj(C_i(a)) s_a C_j(a) (k+Q stack, conduction);
    hole = -x^H (1 (x) Sh) x, electron = x^H (Se (x) 1) x, total = Re(hole + electron)."""
    s = ex.sys
    if any(n % 2 for n in s.orb_per_atom):
        raise ValueError("spinX needs pairs of spin orbitals on every atom")
    sz = np.concatenate([np.tile([0.5, -0.5], n // 2) for n in s.orb_per_atom])
    V = ex.eigvec_k[:, :ex.nv, :]
    C = ex.eigvec_kq[:, ex.nv:, :]
    X = np.asarray(x).reshape(ex.nk, ex.nc, ex.nv)
    Sh = np.einsum("ajk,a,aik->kij", V.conj(), sz, V)
    Se = np.einsum("aik,a,ajk->kij", C.conj(), sz, C)
    hole = -np.einsum("kci,kij,kcj->", X.conj(), Sh, X)
    electron = np.einsum("kiv,kij,kjv->", X.conj(), Se, X)
    return np.array([(hole + electron).real, hole, electron])


def real_space_wavefunction(ex, x, electron_index, hole_index, e_cell, h_cell):
    """ResultTB::realSpaceWavefunction (src/ResultTB.cpp:310-368, 470-482): with
    R = e_cell - h_cell, sum over orbitals alpha of the electron atom and beta of the hole
    atom of |sum_{kcv} C_kc(alpha) conj(V_kv(beta)) e^{i k.R} x_kcv|^2."""
    s = ex.sys
    off = np.cumsum([0] + list(s.orb_per_atom))
    ph = np.exp(1j * (s.kpoints @ (np.asarray(e_cell, float) - np.asarray(h_cell, float))))
    X = (np.asarray(x).reshape(ex.nk, ex.nc * ex.nv) * ph[:, None]).reshape(ex.nk, ex.nc, ex.nv)
    C = ex.eigvec_kq[off[electron_index]:off[electron_index + 1], ex.nv:, :]
    V = ex.eigvec_k[off[hole_index]:off[hole_index + 1], :ex.nv, :]
    amp = np.einsum("ack,bvk,kcv->ab", C, V.conj(), X)
    return float(np.sum(np.abs(amp) ** 2))


# ------------------------------------------------------------------ CRYSTAL outputs

def parse_crystal(path, ncells=20):
    """CRYSTALConfiguration (src/CRYSTALConfiguration.cpp:18-464): lattice (dimension
    lowered for (0, 0, 500) / (0, 500, 0) or zero vectors), motif with species in order of
    first appearance, orbitals per species from the basis-set section, and the overlap and
    Fock matrices (printed lower triangles, column blocks) of the cells with index <= ncells.
    Returns the model dictionary System() takes, plus the raw counts."""
    lines = Path(path).read_text().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    pos = 0
    st = dict(natoms=0, norbitals=0, nsh=0, electrons=0, core=0, ndim=0)
    bravais, motif, shells, orbs = None, None, [], []
    cells, S, F = [], [], []

    def numbers_after(line, key, typ=int):
        rest = line[line.index(key) + len(key):]
        tok = rest.split()
        return typ(tok[0])

    def cell_header(line, key):
        rest = line[line.index(key) + len(key):]
        idx, rest = rest.split("(", 1)
        x, y, z = rest.split(")")[0].split()
        return int(idx), (int(x), int(y), int(z))

    def matrix(n):
        nonlocal pos
        m = np.zeros((n, n))
        cols, row, col = [], 0, 0
        while pos < len(lines):
            line = lines[pos]; pos += 1
            if line == "":
                cols = [int(t) for t in lines[pos].split()]; pos += 1
                line = lines[pos]; pos += 1
                if line == "":
                    continue
            tok = line.split()
            row = int(tok[0])
            for i, t in enumerate(tok[1:]):
                col = cols[i]
                m[row - 1, col - 1] = float(t)
            if row == n and col == n:
                return m
        return m

    def count_numbers(line):
        n = 0
        for t in re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|\S+", line):
            try:
                float(t)
            except ValueError:
                break
            n += 1
        return n

    while pos < len(lines):
        line = lines[pos]; pos += 1
        if "DIRECT LATTICE VECTOR COMPONENTS (ANGSTROM)" in line:
            vec = np.array([[float(x) for x in lines[pos + i].split()[:3]] for i in range(3)])
            pos += 3
            rows = 3
            if np.all(np.abs(vec[2] - [0, 0, 500]) <= 0.1) or np.all(np.abs(vec[2]) <= 0.1):
                rows = 2
                if np.all(np.abs(vec[1] - [0, 500, 0]) <= 0.1) or np.all(np.abs(vec[1]) <= 0.1):
                    rows = 1
            bravais, st["ndim"] = vec[:rows], rows
        elif "N. OF ATOMS PER CELL" in line:
            st["natoms"] = numbers_after(line, "N. OF ATOMS PER CELL")
        elif "NUMBER OF SHELLS" in line:
            st["nsh"] = numbers_after(line, "NUMBER OF SHELLS")
        elif "NUMBER OF AO" in line:
            st["norbitals"] = numbers_after(line, "NUMBER OF AO")
        elif "N. OF ELECTRONS PER CELL" in line:
            st["electrons"] = numbers_after(line, "N. OF ELECTRONS PER CELL")
        elif "CORE ELECTRONS PER CELL" in line:
            st["core"] = numbers_after(line, "CORE ELECTRONS PER CELL")
        elif "ATOM" in line and "SHELL" in line:
            pos += 1
            species, motif, shells = [], [], []
            for _ in range(st["natoms"]):
                tok = lines[pos].split(); pos += 1
                if tok[2] not in species:
                    species.append(tok[2]); shells.append(int(tok[3]))
                motif.append([float(tok[4]), float(tok[5]), float(tok[6]), species.index(tok[2])])
            motif = np.array(motif)
        elif "LOCAL ATOMIC FUNCTIONS BASIS SET" in line:
            pos += 3
            seen, total, ns, last = [], 0, 0, 0
            for a in range(st["natoms"]):
                tok = lines[pos].split(); pos += 1
                if tok[1] in seen:
                    total += orbs[int(motif[a, 3])]
                    continue
                seen.append(tok[1])
                for _ in range(shells[ns]):
                    head = lines[pos].replace("-", " - ", 1).split(); pos += 1
                    last = int(head[2]) if head[1] == "-" else int(head[0])
                    while pos < len(lines) and count_numbers(lines[pos]) == 4:
                        pos += 1
                orbs.append(last - total)
                total = last
                ns += 1
        elif "OVERLAP MATRIX" in line:
            idx, coef = cell_header(line, "OVERLAP MATRIX - CELL N.")
            if idx <= ncells:
                cells.append(sum(bravais[i] * coef[i] for i in range(st["ndim"])))
                S.append(matrix(st["norbitals"]))
        if "FOCK MATRIX" in line:
            idx, coef = cell_header(line, "FOCK MATRIX - CELL N.")
            if idx <= ncells:
                F.append(matrix(st["norbitals"]))
    if st["ndim"] == 0:
        raise ValueError("Finite systems are not yet implemented")
    model = dict(ndim=st["ndim"], bravais=bravais, motif=motif, norbitals=orbs, cells=np.array(cells),
                 H=np.array(F, complex), S=np.array(S, complex), filling=st["electrons"] // 2)
    return model, st
