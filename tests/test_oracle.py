"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

The reference ships no golden vectors for this path (its only hot-path test,
pyparm/tests.py:284-317, checks invariants), so every oracle function is
pinned here by what the mathematics fixes: exact rational arithmetic
(fractions), 50-digit mpmath, closed-form trajectories, brute force and
invariants.  Each check is chosen so that a plausible slip (a dropped /2, a
wrong sign or power, a transposed operand, r instead of r^2) fails it.
"""
import math
import random
from fractions import Fraction as Fr

import mpmath
import numpy as np
import pytest

from oracle import oracle as O

mpmath.mp.dps = 50


# --------------------------------------------------------------- OriginBox
def _remainder_exact(dx: float, L: float) -> float:
    """IEEE remainder from exact rationals: dx - n L, n = nearest, ties even."""
    q = Fr(dx) / Fr(L)
    n = math.floor(q)
    frac = q - n
    if frac > Fr(1, 2) or (frac == Fr(1, 2) and n % 2 == 1):
        n += 1
    r = Fr(dx) - n * Fr(L)
    return float(r)


def test_vec_mod_matches_exact_remainder():
    rng = random.Random(7)
    cases = []
    for _ in range(3000):
        L = rng.uniform(0.5, 200.0)
        cases.append((rng.uniform(-5 * L, 5 * L), L))
    # ties |dx| = (k + 1/2) L and exact multiples of L
    for L in (1.0, 10.7722, 109.4385, 3.0):
        for k in range(-4, 5):
            cases.append(((k + 0.5) * L, L))
            cases.append((k * L, L))
    for dx, L in cases:
        r = O.vec_mod(dx, L)
        assert r == _remainder_exact(dx, L), (dx, L)
        assert abs(r) <= L / 2


def test_vec_mod_tie_rounds_to_even():
    assert O.vec_mod(0.5, 1.0) == 0.5
    assert O.vec_mod(1.5, 1.0) == -0.5
    assert O.vec_mod(-0.5, 1.0) == -0.5
    assert O.vec_mod(2.5, 1.0) == 0.5


# --------------------------------------------------------------- LJ pair
def _lj_table(eps=1.0, sig=1.0, cut=2.5):
    return O.PairTable(O.LJ_ATTRACT_REPULSE, [[eps]], [sig], [cut], [0])


def _lj_exact(rij, eps, sig, cut):
    """V = eps (1 - (sig/r)^6)^2 - eps (1 - cut^-6)^2, F = -dV/dr r_hat,
    in exact rationals (rational in r^2 and r_ij)."""
    d2 = sum(Fr(c) ** 2 for c in rij)
    rsq = d2 / Fr(sig) ** 2
    if rsq > Fr(cut) ** 2:
        return Fr(0), [Fr(0)] * len(rij)
    s = 1 / rsq ** 3
    Vc = Fr(eps) * (1 - 1 / Fr(cut) ** 6) ** 2
    V = Fr(eps) * (1 - s) ** 2 - Vc
    # dV/d(r^2) = eps * 2 (1 - s) * 3 s / r^2 ; F = -2 dV/d(r^2) r_ij
    fs = 12 * Fr(eps) * s * (s - 1) / d2
    return V, [fs * Fr(c) for c in rij]


@pytest.mark.parametrize("seed", range(5))
def test_lj_pair_matches_exact_rationals(seed):
    rng = random.Random(seed)
    L = [1e3] * 3
    for _ in range(200):
        sig = rng.choice([1.0, 1.2, 0.8])
        eps = rng.choice([1.0, 0.5, 2.0])
        r = rng.uniform(0.85 * sig, 2.6 * sig)
        u = np.array([rng.gauss(0, 1) for _ in range(3)])
        rij = u / np.linalg.norm(u) * r
        E, f = O.pair(3, O.PairTable(O.LJ_ATTRACT_REPULSE, [[eps]], [sig], [2.5], [0]),
                      0, 0, rij, np.zeros(3), L)
        Ve, fe = _lj_exact(rij, eps, sig, 2.5)
        scale = abs(float(Ve)) + 1e-300
        assert abs(E - float(Ve)) <= 1e-14 * max(scale, eps)  # cancellation: V = eps(..)^2 - cut_energy
        fn = max(abs(float(c)) for c in fe) + 1e-300
        for k in range(3):
            assert abs(f[k] - float(fe[k])) <= 1e-13 * fn


def test_lj_special_points():
    t = _lj_table()
    L = [100.0] * 3
    # r = sigma: force zero, energy = -cut_energy
    E, f = O.pair(3, t, 0, 0, [1.0, 0, 0], [0, 0, 0], L)
    cutE = (1 - 2.5 ** -6) ** 2
    assert np.all(f == 0) and abs(E + cutE) < 1e-15
    # r = cutoff: energy zero (energy-shifted), force non-zero (not force-shifted)
    E, f = O.pair(3, t, 0, 0, [2.5, 0, 0], [0, 0, 0], L)
    assert abs(E) < 1e-15 and f[0] < 0
    # beyond the cutoff: nothing
    E, f = O.pair(3, t, 0, 0, [2.5000001, 0, 0], [0, 0, 0], L)
    assert E == 0 and np.all(f == 0)
    # repulsive inside sigma: force on atom1 points from atom2 to atom1
    E, f = O.pair(3, t, 0, 0, [0.9, 0, 0], [0, 0, 0], L)
    assert f[0] > 0 and E + cutE > 0


def test_lj_force_is_minus_energy_gradient():
    t = _lj_table(eps=1.3, sig=1.1)
    L = [100.0] * 3
    rng = np.random.default_rng(3)
    for _ in range(50):
        r1 = rng.uniform(-1, 1, 3)
        r1 *= rng.uniform(1.0, 2.6) / np.linalg.norm(r1)
        E, f = O.pair(3, t, 0, 0, r1, [0, 0, 0], L)
        for k in range(3):
            h = 1e-6
            rp, rm = r1.copy(), r1.copy()
            rp[k] += h
            rm[k] -= h
            Ep, _ = O.pair(3, t, 0, 0, rp, [0, 0, 0], L)
            Em, _ = O.pair(3, t, 0, 0, rm, [0, 0, 0], L)
            assert abs(-(Ep - Em) / (2 * h) - f[k]) < 1e-6 * (1 + abs(f[k]))


def test_lj_mixing_rules_and_repulsive_only():
    # two types: eps row lookup (atom1 row, atom2 column), sigma arithmetic mean,
    # cutoff = max(sigcut); eps <= 0 -> purely repulsive WCA-like, cut = 1 sigma
    eps = [[1.0, 0.7], [0.7, -2.0]]
    t = O.PairTable(O.LJ_ATTRACT_REPULSE, eps, [1.0, 1.4], [2.5, 2.0], [0, 1])
    L = [100.0] * 3
    rij = np.array([1.5, 0.3, -0.2])
    E, f = O.pair(3, t, 0, 1, rij, [0, 0, 0], L)
    Ve, fe = _lj_exact(rij, 0.7, 1.2, 2.5)
    assert abs(E - float(Ve)) < 1e-14 and np.allclose(f, [float(c) for c in fe], rtol=1e-13)
    # type 1 - type 1: eps = -2 -> |eps| = 2, cut = 1 (in sigma units), no shift
    rij = np.array([1.3, 0.0, 0.0])
    E, f = O.pair(3, t, 1, 1, rij, [0, 0, 0], L)
    d2 = Fr(1.3) ** 2
    s = 1 / (d2 / Fr(1.4) ** 2) ** 3
    assert abs(E - float(2 * (1 - s) ** 2)) < 1e-14
    E, f = O.pair(3, t, 1, 1, [1.41, 0, 0], [0, 0, 0], L)
    assert E == 0 and f[0] == 0


# ---------------------------------------------------------- Repulsion pair
def test_repulsion_harmonic_exact():
    # n = 2: E = eps (1 - R/sig)^2 / 2, F = eps (1 - R/sig) / sig r_hat;
    # rational R via Pythagorean directions (2,3,6)/7 and (3,4)/5
    t = O.PairTable(O.REPULSION, [1.2, 0.8], [1.0, 1.4], [2.0, 2.0], [0, 1])
    L = [50.0] * 3
    for R in (Fr(1, 2), Fr(9, 10), Fr(11, 10), Fr(6, 5)):
        rij = [R * 2 / 7, R * 3 / 7, -R * 6 / 7]
        E, f = O.pair(3, t, 0, 1, [float(c) for c in rij], [0, 0, 0], L)
        eps = mpmath.sqrt(mpmath.mpf(1.2) * mpmath.mpf(0.8))
        sig = Fr(12, 10)
        if R >= sig:
            assert E == 0 and np.all(f == 0)
            continue
        x = 1 - R / sig
        xm = mpmath.mpf(x.numerator) / x.denominator
        Ee = eps * xm ** 2 / 2
        assert abs(E - float(Ee)) <= 1e-15 * float(abs(Ee)) + 1e-17
        fmag = eps * xm / (mpmath.mpf(sig.numerator) / sig.denominator)
        for k in range(3):
            fe = fmag * mpmath.mpf(rij[k].numerator) / rij[k].denominator / (mpmath.mpf(R.numerator) / R.denominator)
            assert abs(f[k] - float(fe)) <= 2e-15 * float(fmag)


def test_repulsion_general_exponent_gradient_mp():
    # n = (2.5 + 3.0) / 2 = 2.75 : F = -dE/dR r_hat checked against 50-digit mpmath
    t = O.PairTable(O.REPULSION, [1.0, 2.0], [1.0, 1.2], [2.5, 3.0], [0, 1])
    L = [50.0] * 3
    eps = mpmath.sqrt(2)
    sig = mpmath.mpf(1.1)
    n = mpmath.mpf(2.75)
    for R in (0.3, 0.7, 1.0, 1.09):
        E, f = O.pair(3, t, 0, 1, [R, 0, 0], [0, 0, 0], L)
        Rm = mpmath.mpf(R)
        Ee = eps * (1 - Rm / sig) ** n / n
        Fe = -mpmath.diff(lambda r: eps * (1 - r / sig) ** n / n, Rm)
        # conditioning of (1 - R/sig)^n: relative error ~ n ulp / (1 - R/sig)
        cond = 4 * float(n) / float(1 - Rm / sig)
        assert abs(E - float(Ee)) <= 2.2e-16 * cond * float(Ee)
        assert abs(f[0] - float(Fe)) <= 2.2e-16 * cond * float(Fe)


def test_repulsion_params_symmetric_in_atoms():
    t = O.PairTable(O.REPULSION, [1.0, 3.0], [1.0, 1.4], [2.0, 2.5], [0, 1])
    L = [50.0] * 3
    r = np.array([0.4, 0.5, 0.6])
    E1, f1 = O.pair(3, t, 0, 1, r, [0, 0, 0], L)
    E2, f2 = O.pair(3, t, 1, 0, [0, 0, 0], r, L)
    assert E1 == E2 and np.all(f1 == -f2)


# --------------------------------------------------------- neighbour list
def _numpy_pairs(x, L, diam, skin, member):
    """Independent pair-set computation (vectorised, different code path)."""
    x = np.asarray(x)
    pos = x[member]
    out = set()
    M = len(member)
    for i in range(M):
        d = pos[i] - pos[:i]
        d = d - L * np.round(d / L)
        r = np.sqrt(np.sum(d * d, axis=1))
        ok = np.nonzero(r < (diam[i] + diam[:i]) / 2 + skin)[0]
        for j in ok:
            out.add((int(member[i]), int(member[j])))
    return out


@pytest.mark.parametrize("ndim,n,Lside,bidisperse,skin", [
    (3, 300, 7.0, False, 0.4),
    (3, 400, 9.0, True, 0.3),
    (2, 500, 20.0, True, 0.2),
    (3, 60, 3.5, True, 0.4),     # fewer than 3 cells per side
    (2, 50, 2.0, False, 0.1),    # box smaller than the list radius
])
def test_nlist_brute_cells_numpy_agree(ndim, n, Lside, bidisperse, skin):
    rng = np.random.default_rng(n)
    L = np.full(ndim, Lside)
    L[0] *= 1.1  # rectangular
    x = rng.uniform(-3 * Lside, 3 * Lside, (n, ndim))  # unwrapped positions
    diam = np.where(rng.random(n) < 0.5, 1.0, 1.4) if bidisperse else np.full(n, 1.0)
    member = np.arange(n)
    fb, lb = O.nlist(x, L, diam, skin, brute=True)
    fc, lc = O.nlist(x, L, diam, skin, brute=False)
    assert np.array_equal(fb, fc) and np.array_equal(lb, lc)
    # reference order: first ascending, last ascending within first, last < first
    assert np.all(lb < fb)
    key = fb.astype(np.int64) * n + lb
    assert np.all(np.diff(key) > 0)
    if Lside > 2 * (1.4 + skin):  # minimum image unique: numpy round == remainder
        assert set(zip(fb.tolist(), lb.tolist())) == _numpy_pairs(x, L, diam, skin, member)


def test_nlist_member_subset_and_add_order():
    rng = np.random.default_rng(11)
    n = 200
    L = np.full(3, 6.0)
    x = rng.uniform(0, 6.0, (n, 3))
    member = rng.permutation(n)[:150]  # add order is not index order
    diam = rng.uniform(0.8, 1.5, 150)
    fb, lb = O.nlist(x, L, diam, 0.3, member=member, brute=True)
    fc, lc = O.nlist(x, L, diam, 0.3, member=member, brute=False)
    assert np.array_equal(fb, fc) and np.array_equal(lb, lc)
    rank = {int(a): k for k, a in enumerate(member)}
    ranks = [(rank[int(a)], rank[int(b)]) for a, b in zip(fb, lb)]
    assert all(r1 > r2 for r1, r2 in ranks)
    assert ranks == sorted(ranks)


def test_nlist_predicate_is_strict():
    # a pair exactly at the list radius is excluded (|r| < diam + skin)
    x = np.array([[0.0, 0, 0], [1.5, 0, 0], [0, 1.4999999999999998, 0]])
    f, l = O.nlist(x, [50.0] * 3, [1.0, 1.0, 1.0], 0.5, brute=True)
    pairs = set(zip(f.tolist(), l.tolist()))
    assert (1, 0) not in pairs and (2, 0) in pairs


# --------------------------------------------------------------- skin test
def test_skin_check_is_top2_sum():
    rng = np.random.default_rng(5)
    for trial in range(300):
        n = int(rng.integers(1, 40))
        last = rng.normal(size=(n, 3))
        x = last + rng.normal(scale=0.1, size=(n, 3))
        if trial % 3 == 0 and n > 2:  # ties
            x[1] = last[1] + (x[0] - last[0])
        d = np.sqrt(np.sum((x - last) ** 2, axis=1))
        top = np.sort(d)[::-1]
        s = top[0] + (top[1] if n > 1 else 0.0)
        for skin in (s, np.nextafter(s, 0), np.nextafter(s, 10), 0.5 * s, 2 * s):
            assert O.skin_check(x, last, skin) == bool(s >= skin)


# ---------------------------------------------------------- forces, sums
def _random_liquid(n, rho, ndim, seed, bidisperse=False):
    rng = np.random.default_rng(seed)
    Lside = (n / rho) ** (1 / ndim)
    # jittered lattice keeps the minimum separation reasonable
    m = int(math.ceil(n ** (1 / ndim)))
    g = np.stack(np.meshgrid(*[np.arange(m)] * ndim, indexing="ij"), -1).reshape(-1, ndim)[:n]
    x = (g + 0.5) * (Lside / m) + rng.normal(scale=0.05, size=(n, ndim))
    types = (rng.random(n) < 0.5).astype(np.int32) if bidisperse else np.zeros(n, np.int32)
    return x, np.full(ndim, Lside), types


def test_forces_newton3_virial_stress_fd():
    x, L, types = _random_liquid(216, 0.8, 3, 1)
    t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], types)
    first, last = O.nlist(x, L, np.full(216, 2.5), 0.3)
    r = O.forces(x, L, t, first, last)
    f = r["f"]
    assert np.abs(f.sum(0)).max() < 1e-11 * np.abs(f).max()
    assert abs(np.trace(r["S"]) - r["W"]) < 1e-11 * abs(r["W"])
    # finite-difference gradient of U for a few atoms (list radius covers moves)
    for i in (0, 17, 100):
        for k in range(3):
            h = 1e-6
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            Up = O.forces(xp, L, t, first, last)["U"]
            Um = O.forces(xm, L, t, first, last)["U"]
            assert abs(-(Up - Um) / (2 * h) - f[i, k]) < 1e-5 * (1 + abs(f[i, k]))


def test_virial_equals_sum_r_dot_f_for_isolated_cluster():
    # no wrapping in a huge box: sum_pairs r_ij . f_ij == sum_i x_i . f_i
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 4, (40, 3))
    L = np.full(3, 1e4)
    t = O.PairTable(O.REPULSION, [1.0], [1.3], [2.0], np.zeros(40, np.int32))
    first, last = O.nlist(x, L, np.full(40, 1.3), 0.2)
    r = O.forces(x, L, t, first, last)
    assert abs(r["W"] - np.sum(x * r["f"])) < 1e-12 * (abs(r["W"]) + 1)
    assert r["overlaps"] == r["contacts"] > 0


def test_energy_is_sum_of_pair_energies():
    x, L, types = _random_liquid(100, 0.9, 2, 4, bidisperse=True)
    t = O.PairTable(O.REPULSION, [1.0, 1.0], [1.0, 1.4], [2.0, 2.0], types)
    diam = np.where(types == 1, 1.4, 1.0)
    first, last = O.nlist(x, L, diam, 0.1)
    r = O.forces(x, L, t, first, last)
    Es = sum(O.pair(2, t, types[a], types[b], x[a], x[b], L)[0] for a, b in zip(first, last))
    assert abs(r["U"] - Es) < 1e-12 * (abs(Es) + 1e-300)


# ------------------------------------------------------------ integrators
def _two_atom_system(R0, sig=1.0, eps=1.0, m=1.0, collection="verlet"):
    t = O.PairTable(O.REPULSION, [eps], [sig], [2.0], [0, 0])
    s = O.System([m, m], [20.0] * 3, t, [sig, sig], 0.3)
    x = np.array([[0.0, 0, 0], [R0, 0, 0]])
    s.set_state(x=x, v=np.zeros((2, 3)), a=np.zeros((2, 3)))
    s.initialize()
    return s


def test_verlet_harmonic_contact_closed_form():
    # relative coordinate u = R - sigma obeys u'' = -w^2 u while in contact,
    # w^2 = 2 eps / (m sigma^2); velocity Verlet gives u_n = u_0 cos(n theta),
    # cos(theta) = 1 - (w dt)^2 / 2, exactly (up to rounding).
    sig, eps, m, dt = 1.0, 1.0, 1.0, 0.01
    u0 = -0.2
    s = _two_atom_system(sig + u0, sig, eps, m)
    w2 = 2 * eps / (m * sig * sig)
    theta = math.acos(1 - w2 * dt * dt / 2)
    n = 1
    while (n + 1) * theta < math.pi / 2 * 0.95:
        s.verlet(dt, 1)
        x = s.state()[0]
        u = (x[1, 0] - x[0, 0]) - sig
        assert abs(u - u0 * math.cos(n * theta)) < 1e-13, n
        n += 1
    assert n > 50


def test_verlet_time_reversible():
    x, L, types = _random_liquid(64, 0.7, 3, 9)
    t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], types)
    s = O.System(np.ones(64), L, t, np.full(64, 2.5), 0.4)
    rng = np.random.default_rng(0)
    v0 = rng.normal(size=(64, 3))
    s.set_state(x=x, v=v0, a=np.zeros((64, 3)))
    s.initialize()
    x0 = s.state()[0].copy()
    s.verlet(0.002, 200)
    xs, vs, _, _ = s.state()
    s.set_state(v=-vs)
    s.verlet(0.002, 200)
    x1, v1, _, _ = s.state()
    assert np.abs(x1 - x0).max() < 1e-9
    assert np.abs(-v1 - v0).max() < 1e-8


def test_fixed_atoms_do_not_move():
    x, L, types = _random_liquid(27, 0.6, 3, 3)
    masses = np.ones(27)
    masses[5] = np.inf
    masses[6] = 0.0
    t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], types)
    s = O.System(masses, L, t, np.full(27, 2.5), 0.4)
    s.set_state(x=x, v=np.ones((27, 3)))
    s.initialize()
    s.verlet(0.001, 10)
    xs, vs, a, _ = s.state()
    assert np.all(xs[5] == x[5]) and np.all(xs[6] == x[6])
    assert np.all(vs[5] == 0) and np.all(a[6] == 0)


def test_damped_constants_allen_tildesley():
    c0, c1, c2 = O.damped_constants(0.01, 0.5)
    g = mpmath.mpf(0.5) * mpmath.mpf(0.01)
    assert abs(c0 - float(mpmath.exp(-g))) < 1e-16
    assert abs(c1 - float((1 - mpmath.exp(-g)) / g)) < 1e-15
    assert abs(c2 - float((1 - (1 - mpmath.exp(-g)) / g) / g)) < 1e-13
    assert O.damped_constants(0.01, 0.0) == (1.0, 1.0, 0.5)


def test_damped_free_particle_closed_form():
    # no contacts: x_n = x0 + v0 (1 - exp(-g n dt)) / g, v_n = v0 exp(-g n dt)
    t = O.PairTable(O.REPULSION, [1.0], [1.0], [2.0], [0, 0])
    s = O.System([1.0, 1.0], [100.0] * 3, t, [1.0, 1.0], 0.1)
    v0 = np.array([[0.3, -0.2, 0.1], [0.0, 0.5, 0.0]])
    x0 = np.array([[0.0, 0, 0], [50.0, 0, 0]])
    s.set_state(x=x0, v=v0)
    s.initialize()
    g, dt, n = 0.7, 0.01, 300
    s.damped(dt, g, n)
    x, v, _, _ = s.state()
    e = math.exp(-g * n * dt)
    assert np.allclose(v, v0 * e, rtol=1e-12, atol=1e-15)
    assert np.allclose(x, x0 + v0 * (1 - e) / g, rtol=1e-11, atol=1e-13)


def test_damped_without_damping_is_verlet_bitwise():
    """CollectionDamped with damping <= 0 has c0 = c1 = 1, c2 = 1/2
    (src/collection.cpp:342-348), and then every update of its step
    (:356-381) rounds exactly as CollectionVerlet's (:442-469): (0.5 dt) dt
    = (dt dt)/2 and v*1 = v exactly.  The two trajectories of an LJ liquid
    (with fixed atoms and rebuilds) must agree bit for bit; this pins the
    acceleration terms a c2 dt^2, a dt (c1 - c2) and a dt c2."""
    n = 64
    x, L, types = _random_liquid(n, 0.8, 3, 11)
    masses = np.ones(n)
    masses[3] = np.inf
    rng = np.random.default_rng(5)
    v = rng.normal(size=(n, 3))
    t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], types)
    runs = []
    for kind in ("verlet", "damped"):
        s = O.System(masses, L, t, np.full(n, 2.5), 0.3)
        s.set_state(x=x, v=v, a=np.zeros_like(x))
        s.initialize()
        for _ in range(60):
            if kind == "verlet":
                s.verlet(0.003, 1)
            else:
                s.damped(0.003, 0.0, 1)
        runs.append((s.state(), s.which()))
    (sa, wa), (sb, wb) = runs
    assert wa == wb and wa > 1
    for qa, qb in zip(sa, sb):
        assert np.array_equal(qa, qb)


def _damped_pair_run(dt, T, gamma, sig, u0):
    """Two unit-mass atoms in harmonic contact (RepulsionPair exponent 2,
    eps = 1) on the x axis, released from rest at separation sig + u0 < sig;
    CollectionDamped with damping gamma up to time T.  Returns the
    separation at T."""
    t = O.PairTable(O.REPULSION, [1.0], [sig], [2.0], [0, 0])
    s = O.System([1.0, 1.0], [50.0] * 3, t, [sig, sig], 0.2)
    R0 = sig + u0
    s.set_state(x=np.array([[R0 / 2, 0, 0], [-R0 / 2, 0, 0]]), v=np.zeros((2, 3)),
                a=np.zeros((2, 3)))
    s.initialize()
    s.damped(dt, gamma, int(round(T / dt)))
    x, _, _, _ = s.state()
    return x[0, 0] - x[1, 0]


def test_damped_harmonic_contact_converges_to_analytic_oscillator():
    """The separation R of a harmonic contact obeys u'' + gamma u' + w^2 u = 0
    with u = R - sig and w^2 = 2 eps / (sig^2 m) (force eps (1 - R/sig)/sig on
    each atom, relative acceleration twice that; each atom's velocity damped
    at rate gamma).  Overdamped from rest the contact never opens, so the
    closed form holds for all t.  The damped Verlet step is second order: the
    error at fixed T drops ~4x per halving of dt.  A c2 <-> c1 slip makes it
    first order (or inconsistent) and fails the ratio."""
    sig, gamma, u0, T = 1.0, 3.0, -0.2, 2.0
    w2 = 2.0 / (sig * sig)
    disc = math.sqrt(gamma * gamma - 4 * w2)
    r1, r2 = (-gamma + disc) / 2, (-gamma - disc) / 2
    # u(0) = u0, u'(0) = 0
    A = u0 * r2 / (r2 - r1)
    B = -u0 * r1 / (r2 - r1)
    exact = sig + A * math.exp(r1 * T) + B * math.exp(r2 * T)
    errs = [abs(_damped_pair_run(dt, T, gamma, sig, u0) - exact)
            for dt in (0.02, 0.01, 0.005)]
    assert errs[0] < 1e-3
    for e_big, e_small in zip(errs, errs[1:]):
        assert 3.3 < e_big / e_small < 4.7, errs


def test_observables_closed_form():
    t = O.PairTable(O.REPULSION, [1.0], [1.0], [2.0], np.zeros(4, np.int32))
    masses = [1.0, 2.0, np.inf, 0.5]
    s = O.System(masses, [100.0] * 3, t, [1.0] * 4, 0.1)
    x = np.array([[0, 0, 0], [10, 0, 0], [20, 0, 0], [30, 0, 0]], float)
    v = np.array([[1, 0, 0], [0, 2, 0], [5, 5, 5], [0, 0, -4]], float)
    s.set_state(x=x, v=v)
    K = 0.5 * 1 * 1 + 0.5 * 2 * 4 + 0.5 * 0.5 * 16  # infinite mass skipped
    assert abs(s.kinetic_energy() - K) < 1e-15
    p = np.array([1.0, 4.0, -2.0])
    assert np.allclose(s.momentum(), p, rtol=0, atol=1e-15)
    assert s.mass() == 3.5
    vc = p / 3.5
    Kc = sum(0.5 * m * np.dot(vi - vc, vi - vc) for m, vi in zip([1, 2, 0.5], v[[0, 1, 3]]))
    assert abs(s.temp(True) - 2 * Kc / (9 - 3)) < 1e-14
    assert abs(s.temp(False) - 2 * K / 9) < 1e-14


@pytest.mark.slow
def test_random_hertzian_verlet_invariants():
    """Mirrors the assertions of RandomHertzianVerletTest (pyparm/tests.py:
    284-317): N=12 bidisperse harmonic spheres, phi=0.3, dt=0.01; per-step
    |dE/E| < 1e-2, mean T within 10 % of 1, std(E)/mean(E) < 1e-2."""
    N = 12
    sig = np.array([1.0] * 6 + [1.4] * 6)
    masses = sig ** 3
    Vs = np.sum((sig / 2) ** 3) * 4 / 3 * np.pi
    Lside = (Vs / 0.3) ** (1 / 3)
    types = np.array([0] * 6 + [1] * 6, np.int32)
    t = O.PairTable(O.REPULSION, [1.2, 1.2], [1.0, 1.4], [2.0, 2.0], types)
    s = O.System(masses, [Lside] * 3, t, sig, 0.4)
    rng = np.random.default_rng(131)
    x = rng.uniform(0.0, Lside, (N, 3))
    v = rng.normal(size=(N, 3)) / masses[:, None]
    s.set_state(x=x, v=v, f=rng.normal(size=(N, 3)))
    s.update_list(True)
    s.initialize()

    def scale_to_temp(T):
        tt = s.temp(True)
        _, vv, _, _ = s.state()
        s.set_state(v=vv * math.sqrt(T / tt))

    scale_to_temp(1.0)
    for _ in range(1000):
        s.verlet(0.01, 1)
        scale_to_temp(1.0)
    E_last = s.energy()
    Es, Ts = [], []
    for _ in range(300):
        for _ in range(10):
            s.verlet(0.01, 1)
            E = s.energy()
            assert abs(E - E_last) <= 1e-2 * abs(E_last) + 1e-8
            E_last = E
        Es.append(E)
        Ts.append(s.temp(True))
    assert abs(np.mean(Ts) - 1.0) < 0.1 * 1.0 + 0.15  # short run: looser window
    assert np.std(Es) / np.mean(Es) < 1e-2
