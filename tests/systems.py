"""Seeded test systems shared by the CPU and GPU tests (test infrastructure)."""
from __future__ import annotations

import math

import numpy as np


def jittered_lattice(n, rho, ndim, seed, jitter=0.08, offset=None):
    """n points at number density rho: a simple lattice with Gaussian jitter,
    shifted by a random offset so positions are not box-aligned, and NOT
    wrapped (unwrapped coordinates are what the reference keeps)."""
    rng = np.random.default_rng(seed)
    L = (n / rho) ** (1.0 / ndim)
    m = int(math.ceil(n ** (1.0 / ndim) - 1e-9))
    g = np.stack(np.meshgrid(*[np.arange(m)] * ndim, indexing="ij"), -1).reshape(-1, ndim)
    g = g[rng.permutation(len(g))[:n]]
    x = (g + 0.5) * (L / m) + rng.normal(scale=jitter, size=(n, ndim))
    x += rng.uniform(-L, L, ndim) if offset is None else offset
    # a few atoms displaced by whole box lengths: positions are unwrapped
    k = rng.choice(n, size=max(1, n // 50), replace=False)
    x[k] += rng.integers(-3, 4, size=(len(k), ndim)) * L
    return x, np.full(ndim, L)


def lj_params(n, sigma=1.0, eps=1.0, sigcut=2.5):
    """Single-type NListed<IEpsSigCutAtom, LJAttractRepulsePair> table."""
    return dict(eps=[[eps]], sigma=[sigma], third=[sigcut], types=np.zeros(n, np.int32),
                diam=np.full(n, sigma * sigcut))


def bidisperse_repulsion(n, seed, eps=(1.0, 1.0), sigma=(1.0, 1.4), exponent=(2.0, 2.0)):
    rng = np.random.default_rng(seed + 1000)
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: n // 2]] = 1
    sig = np.asarray(sigma)[types]
    return dict(eps=list(eps), sigma=list(sigma), third=list(exponent), types=types,
                diam=sig)


def gaussian_velocities(n, ndim, T, seed, masses=None):
    rng = np.random.default_rng(seed + 2000)
    m = np.ones(n) if masses is None else np.asarray(masses, float)
    v = rng.normal(size=(n, ndim)) * np.sqrt(T / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v


def normalised_force_error(f_dev, f_ref, fabs):
    """max_i |f_dev_i - f_ref_i| / sum_j |f_ij|  (per-atom normaliser; atoms
    with no pair force are compared absolutely)."""
    d = np.sqrt(np.sum((f_dev - f_ref) ** 2, axis=1))
    den = np.where(fabs > 0, fabs, 1.0)
    return float(np.max(d / den)) if len(d) else 0.0
