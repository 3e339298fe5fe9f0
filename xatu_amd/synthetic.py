"""Seeded synthetic tight-binding models used as benchmark workloads.

A model is written in Xatu's ``.model`` text format (the format parsed by
SystemConfiguration), so it enters the pipeline exactly like a user's model.

Default model (18 orbitals): triangular lattice a = 3.19 Angstrom, three atoms at
fractional (0,0), (1/3,2/3), (2/3,1/3), six orbitals per atom. On every atom three
orbitals are valence-like with onsite energies U(-4,-2) eV and three are
conduction-like with U(2,4) eV; the filling is half the orbitals. Hoppings couple
every orbital pair to all cells R = m1 a1 + m2 a2 with |m1|, |m2| <= 3:
t_ij(R) = (N(0, s) + i N(0, s)) / sqrt(2) * exp(-|R|/a) with s = 0.5 eV, and
H(-R) = H(R)^H exactly (the R = 0 block is made Hermitian the same way).
The larger model (36 orbitals) uses four atoms, nine orbitals each, the same
distributions, and the additional atom at fractional (1/2, 0); half of its orbitals are
valence-like as well (five on the first two atoms, four on the others: 18 of 36, the
filling).
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

TRIANGULAR_A = 3.19


def model_arrays(natoms=3, orb_per_atom=6, seed=20240328, a=TRIANGULAR_A, reach=3, sigma=0.5,
                 fractional=None):
    """Return (bravais 2x3, motif natoms x 4, cells ncells x 3, H ncells x norb x norb, filling)."""
    rng = np.random.default_rng(seed)
    if fractional is None:
        fractional = [(0.0, 0.0), (1 / 3, 2 / 3), (2 / 3, 1 / 3), (0.5, 0.0)][:natoms]
    a1 = np.array([a, 0.0, 0.0])
    a2 = np.array([a / 2, a * math.sqrt(3) / 2, 0.0])
    bravais = np.array([a1, a2])
    motif = np.array([[*(f1 * a1 + f2 * a2)[:3], 0.0] for f1, f2 in fractional])
    motif[:, 3] = 0.0
    norb = natoms * orb_per_atom
    # half of all orbitals are valence-like: orb_per_atom // 2 on every atom, and with an
    # odd orb_per_atom one more on the first atoms until the total is norb // 2
    extra = norb // 2 - natoms * (orb_per_atom // 2)
    onsite = []
    for at in range(natoms):
        nval = orb_per_atom // 2 + (1 if at < extra else 0)
        onsite += list(rng.uniform(-4.0, -2.0, nval)) + list(rng.uniform(2.0, 4.0, orb_per_atom - nval))
    onsite = np.array(onsite)

    ms = [(m1, m2) for m2 in range(-reach, reach + 1) for m1 in range(-reach, reach + 1)]
    H = {}
    for m1, m2 in ms:
        if (m1, m2) in H:
            continue
        R = m1 * a1 + m2 * a2
        decay = math.exp(-np.linalg.norm(R) / a)
        T = (rng.normal(0.0, sigma, (norb, norb)) + 1j * rng.normal(0.0, sigma, (norb, norb))) / math.sqrt(2)
        T *= decay
        if (m1, m2) == (0, 0):
            U = np.triu(T, 1)
            H[(0, 0)] = U + U.conj().T + np.diag(onsite)
        else:
            H[(m1, m2)] = T
            H[(-m1, -m2)] = T.conj().T
    cells = np.array([m1 * a1 + m2 * a2 for m1, m2 in ms])
    Hs = np.array([H[m] for m in ms])
    filling = norb // 2
    return bravais, motif, cells, Hs, filling


def model_text(natoms=3, orb_per_atom=6, seed=20240328, **kw) -> str:
    bravais, motif, cells, Hs, filling = model_arrays(natoms, orb_per_atom, seed, **kw)
    out = ["# dimension", "2", "# norbitals", str(orb_per_atom), "# bravaislattice"]
    out += ["  ".join(f"{x:.17g}" for x in v) for v in bravais]
    out += ["# motif"]
    out += ["  ".join(f"{x:.17g}" for x in row[:3]) + "  0" for row in motif]
    out += ["# bravaisvectors"]
    out += ["  ".join(f"{x:.17g}" for x in v) for v in cells]
    out += ["# hamiltonian"]
    for Hm in Hs:
        for row in Hm:
            out.append("  ".join(f"{z.real:.17g} {z.imag:+.17g}j" for z in row))
        out.append("&")
    out += ["# filling", str(filling), "#"]
    return "\n".join(out)


def write_model(path, natoms=3, orb_per_atom=6, seed=20240328, **kw) -> Path:
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(model_text(natoms, orb_per_atom, seed, **kw))
    return p
