"""Shared parity cases: oracle pipelines and C-ABI contexts fed with the same band stacks."""
from __future__ import annotations

from pathlib import Path

import numpy as np

import oracle as O

MODELS = Path(__file__).resolve().parent / "golden" / "models"


def oracle_case(model, ncell, nbands=1, nrmbands=0, params=(1, 1, 10), Q=(0, 0, 0), exchange=False,
                factor=1, shift=None, cutoff=None, scissor=0.0, potential="keldysh",
                exchange_potential="keldysh", bands=None):
    s = O.System(str(model))
    if factor > 1:
        s.reduced_brillouin_zone_mesh(ncell, factor)
    else:
        s.brillouin_zone_mesh(ncell)
    if shift is not None:
        s.shift_bz(shift)
    ex = O.Exciton(s, ncell, nbands, nrmbands, params, Q=Q, bands=bands)
    ex.exchange = exchange
    ex.scissor = scissor
    ex.potential = potential
    ex.exchange_potential = exchange_potential
    if cutoff is not None:
        ex.cutoff = cutoff
    ex.initialize()
    return ex


def k_offsets(ex):
    """Integer offsets of every k-point from k-point 0 in units of 1/(ncell*factor)."""
    s = ex.sys
    n = ex.ncell * s.factor
    dk = s.kpoints - s.kpoints[0]
    coords = (s.inv @ (s.recip @ dk.T)).T  # reciprocal coordinates
    return np.rint(coords * n).astype(np.int32)


def gpu_context(ex, device=0, n_reciprocal=None):
    """Context with the oracle's lattice, band stacks and interaction; n_reciprocal selects
    the reciprocal-space mode with that many reciprocal vectors per axis."""
    from xatu_amd import capi
    s = ex.sys
    ctx = capi.Context(device)
    ctx.set_lattice(s.bravais[:s.ndim], s.motif[:, :3], s.orb_per_atom, ex.ncell * s.factor)
    ctx.set_bands(ex.nv, ex.nc, ex.eigval_k, ex.eigvec_k, k_offsets(ex), ex.eigval_kq, ex.eigvec_kq)
    pot = {"keldysh": capi.XATU_POTENTIAL_KELDYSH, "coulomb": capi.XATU_POTENTIAL_COULOMB}
    ctx.fourier_sums(ex.eps_m, ex.eps_s, ex.r0, ex.cutoff, ex.ncell, ex.total_cells,
                     regularization=ex.regularization, potential=pot[ex.potential],
                     exchange_potential=pot[ex.exchange_potential], exchange=ex.exchange, Q=tuple(ex.Q),
                     mode=0 if n_reciprocal is None else capi.XATU_MODE_RECIPROCAL,
                     n_reciprocal=1 if n_reciprocal is None else n_reciprocal)
    return ctx


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
