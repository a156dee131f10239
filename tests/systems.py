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


def _pair_scales(x, L, table, first, last):
    """Per pair within its interaction range: (term, sens, r).  term is the
    scale of the terms the pair force is formed from, which floating point
    rounds separately: |c1| s^2 / r + |c2| s / r for the LJ forms (f = c1 s^2
    / r + c2 s / r with s = (sigma / r)^6; the two terms cancel near the
    minimum), eps / sigma for the repulsion (whose 1 - R / sigma cancels near
    contact).  sens bounds |f'(r)| + |f(r)| / r, how far the pair force vector
    moves per unit change of the displacement: 14 term / r for the LJ forms,
    term ((n - 1) / sigma + 1 / r) for the repulsion of exponent n >= 2."""
    from oracle import oracle as O

    t1, t2 = table.types[first], table.types[last]
    eps = np.abs(table.eps[t1, t2] if table.eps.ndim == 2
                 else np.sqrt(table.eps[t1] * table.eps[t2]))
    sig = (table.sigma[t1, t2] if table.sigma.ndim == 2
           else (table.sigma[t1] + table.sigma[t2]) / 2)
    d = x[first] - x[last]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(1))
    if table.kind in (O.REPULSION, O.REPULSION_II):
        n = (table.third[t1] + table.third[t2]) / 2
        term = np.where(r < sig, eps / sig, 0.0)
        sens = term * (np.maximum(n - 1.0, 1.0) / sig + 1.0 / r)
    else:
        cut = (1.0 if table.kind == O.LJ_REPULSE
               else np.maximum(table.third[t1], table.third[t2]))
        s6 = (sig / r) ** 6
        if table.kind in (O.LJ_CUT, O.LJ_CUT_II):  # 4 eps (s^2 - s): 48 s^2 - 24 s
            t = 24 * eps * s6 * (2 * s6 + 1) / r
        else:  # eps (1 - s)^2: 12 s^2 - 12 s
            t = 12 * eps * s6 * (s6 + 1) / r
        term = np.where(r < cut * sig, t, 0.0)
        sens = 14.0 * term / r
    return term, sens, r


def pair_term_scale(x, L, table, first, last):
    """Per atom, the sum over its pairs of the scale of the terms each pair
    force is formed from (see _pair_scales).  table is the oracle's
    PairTable."""
    x = np.asarray(x, float)
    n = len(x)
    if len(first) == 0:
        return np.zeros(n)
    term, _, _ = _pair_scales(x, L, table, first, last)
    return np.bincount(first, term, n) + np.bincount(last, term, n)


def assert_fp64_forces(f_dev, ref, x, L, table, first, last):
    """The FP64 bar, per atom: |f_dev - f_ref| <= 1e-12 sum_j |f_ij|
    + 1e-13 T + sum_j sens_ij delta_ij.  T = pair_term_scale: FP64 rounds
    each term of a pair force, not their cancelling sum.  delta_ij bounds the
    difference of the two displacements: the reference forms
    remainder(x_i - x_j, L) from unwrapped positions, so x_i - x_j rounds at
    its own magnitude, and the device differences two offsets each rounded
    at ulp(L): |dr| <= 4 sqrt(D) 2^-52 (|x_i - x_j|_inf + L)."""
    x = np.asarray(x, float)
    n, ndim = x.shape
    d = np.sqrt(((np.asarray(f_dev) - ref["f"]) ** 2).sum(1))
    bar = 1e-12 * ref["fabs"]
    if len(first):
        term, sens, _ = _pair_scales(x, L, table, first, last)
        dlt = 4.0 * np.sqrt(ndim) * 2.0 ** -52 * (
            np.abs(x[first] - x[last]).max(1) + np.max(L))
        g = 1e-13 * term + sens * dlt
        bar = bar + np.bincount(first, g, n) + np.bincount(last, g, n)
    assert np.all(d <= bar), float(np.max(d / np.maximum(bar, 1e-300)))


def normalised_force_error(f_dev, f_ref, fabs):
    """max_i |f_dev_i - f_ref_i| / sum_j |f_ij|  (per-atom normaliser; atoms
    with no pair force are compared absolutely)."""
    d = np.sqrt(np.sum((f_dev - f_ref) ** 2, axis=1))
    den = np.where(fabs > 0, fabs, 1.0)
    return float(np.max(d / den)) if len(d) else 0.0
