"""Seeded synthetic workloads for the benchmark and the large-N tests.

lj_liquid: the Lennard-Jones liquid of the LJatoms.cpp-style runs: N atoms of
mass 1 in a cubic OriginBox at number density rho, positions uniform in
[-L/2, L/2) (OriginBox::rand_loc, src/box.hpp:145-151) then pushed apart on
the device until no two atoms are closer than min_sep (a harmonic Repulsion
of diameter min_sep relaxed with CollectionDamped), velocities Gaussian at
temperature T with zero net momentum, rescaled to exactly T
(Collection::scale_velocities_to_temp, src/collection.cpp:31-35).
"""
from __future__ import annotations

import numpy as np

from . import capi


def box_length(n, rho, ndim=3):
    return (n / rho) ** (1.0 / ndim)


def push_apart(x, L, min_sep, seed_diam=None, max_rounds=400, log=None):
    """Relax overlapping positions on the device until every pair is at least
    min_sep apart (harmonic repulsion of diameter min_sep * (1 + 1e-3), damped
    dynamics)."""
    n, ndim = x.shape
    sig = min_sep * (1.0 + 1e-3) if seed_diam is None else seed_diam
    s = capi.System(ndim, np.ones(n), capi.FP64)
    try:
        s.set_box(L)
        s.nlist_set(0.15 * sig, np.full(n, sig))
        s.set_repulsion([1.0], [sig], [2.0], np.zeros(n, np.int32))
        s.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
        s.initialize()
        # A pair closer than min_sep has energy >= (1 - min_sep/sig)^2 / 2, so
        # a total energy below that bound proves every pair is >= min_sep apart
        # (relaxation towards contact at sig itself is only asymptotic).
        bound = 0.5 * (1.0 - min_sep / sig) ** 2
        for r in range(max_rounds):
            s.damped(0.1 * sig, 2.0, 50)
            U = s.pair_sums()["U"]
            if log and r % 20 == 0:
                log(f"  push_apart round {r}: overlap energy {U:.3e} (bound {bound:.3e})")
            if U < bound:
                break
        else:
            raise RuntimeError("push_apart did not remove all overlaps")
        return s.positions()
    finally:
        s.close()


def gaussian_velocities(n, ndim, T, rng, masses=None):
    m = np.ones(n) if masses is None else np.asarray(masses, float)
    v = rng.normal(size=(n, ndim)) * np.sqrt(T / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    # scale to exactly T with ndof = ndim N - ndim (COM removed)
    K = 0.5 * np.sum(m[:, None] * v * v)
    Tnow = 2 * K / (ndim * n - ndim)
    return v * np.sqrt(T / Tnow)


def lj_liquid(n, rho=0.8, T=1.0, min_sep=0.8, seed=2024, ndim=3, log=None):
    rng = np.random.default_rng(seed)
    L = box_length(n, rho, ndim)
    x = rng.uniform(-L / 2, L / 2, size=(n, ndim))
    x = push_apart(x, np.full(ndim, L), min_sep, log=log)
    v = gaussian_velocities(n, ndim, T, rng)
    return x, v, np.full(ndim, L)


def bidisperse_packing(n, phi, seed, ndim, sigmas=(1.0, 1.4)):
    """Bidisperse discs / spheres, 50/50 (odd n: the extra one small), uniform
    positions in a box at packing fraction phi."""
    rng = np.random.default_rng(seed)
    nl = n // 2
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[:nl]] = 1
    sig = np.asarray(sigmas)[types]
    if ndim == 3:
        vs = np.sum(np.pi / 6 * sig ** 3)
    else:
        vs = np.sum(np.pi / 4 * sig ** 2)
    L = (vs / phi) ** (1.0 / ndim)
    x = rng.uniform(-L / 2, L / 2, size=(n, ndim))
    return x, types, np.full(ndim, L)
