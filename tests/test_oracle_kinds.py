"""Pins the oracle's further pair kinds, its exclusions and its
CollectionNLCG before they are trusted as checkers (CPU only).

  * LennardJonesCutPair (EpsSigCutAtom / IEpsISigCutAtom), LJRepulsePair and
    RepulsionPair over IEpsISigExpAtom: exact rational energies and forces at
    random rational points (they are rational in r^2 and r_ij), the special
    points (F = 0 at r = sigma, V = 0 at the cut, nothing beyond it), and the
    mixing rules, with parameters chosen so a wrong rule gives another value
    (sqrt of the product vs arithmetic mean; row lookup vs mixing);
  * NeighborList::ignore: the pair set equals an independent numpy brute
    force minus the excluded pairs; ignore forces the next rebuild;
  * CollectionNLCG: what the minimiser is for -- from a random packing it
    reaches mechanical equilibrium at the goal pressure (the virial pressure
    equals P0, the forces vanish) and H = U + P0 V falls; FixedL keeps the
    other box edges.
"""
import math
import random
from fractions import Fraction as Fr

import mpmath
import numpy as np
import pytest

from oracle import oracle as O

mpmath.mp.dps = 50


def _lj_cut_exact(rij, eps, sig, cut, shift=True):
    """V = eps (1 - s)^2 - eps (1 - cut^-6)^2 (shift) with s = (sig / r)^6;
    F = 12 eps s (s - 1) / r^2 r_ij; zero beyond r^2 / sig^2 > cut^2."""
    d2 = sum(Fr(c) ** 2 for c in rij)
    rsq = d2 / Fr(sig) ** 2
    if rsq > Fr(cut) ** 2:
        return Fr(0), [Fr(0)] * len(rij)
    s = 1 / rsq ** 3
    V = Fr(eps) * (1 - s) ** 2
    if shift:
        V -= Fr(eps) * (1 - 1 / Fr(cut) ** 6) ** 2
    fs = 12 * Fr(eps) * s * (s - 1) / d2
    return V, [fs * Fr(c) for c in rij]


def _check(E, f, Ve, fe, eps, rij=None, sig=None):
    """Energy within 1e-14 of its scale; each force component within 1e-14
    of the scale of the two terms 12 eps s^2 and 12 eps s (over r), which
    cancel near the minimum r = sigma."""
    assert abs(E - float(Ve)) <= 1e-14 * max(abs(float(Ve)), eps)
    fn = max(abs(float(c)) for c in fe) + 1e-300
    if rij is not None:
        r2 = float(np.dot(rij, rij))
        s = (sig * sig / r2) ** 3
        fn = max(fn, 12 * eps * s * (s + 1) / math.sqrt(r2))
    for k in range(len(f)):
        assert abs(f[k] - float(fe[k])) <= 1e-14 * fn


@pytest.mark.parametrize("seed", range(3))
def test_lj_cut_mixing_and_exact_rationals(seed):
    # EpsSigCutAtom: eps = sqrt(e1 e2) (1 * 0.25 -> 0.5), sigma = (1 + 1.2) / 2,
    # cut = max(2.5, 2.2)
    t = O.PairTable(O.LJ_CUT, [1.0, 0.25], [1.0, 1.2], [2.5, 2.2], [0, 1])
    rng = random.Random(seed)
    L = [1e3] * 3
    for _ in range(150):
        r = rng.uniform(0.95, 2.8)
        u = np.array([rng.gauss(0, 1) for _ in range(3)])
        rij = u / np.linalg.norm(u) * r
        E, f = O.pair(3, t, 0, 1, rij, np.zeros(3), L)
        Ve, fe = _lj_cut_exact(rij, 0.5, 1.1, 2.5)
        _check(E, f, Ve, fe, 0.5, rij, 1.1)
    # same-type pairs use their own sigma and cut
    E, f = O.pair(3, t, 1, 1, [1.2, 0, 0], [0, 0, 0], L)
    assert f[0] == 0 and abs(E + 0.25 * (1 - 2.2 ** -6) ** 2) < 1e-15
    E, f = O.pair(3, t, 1, 1, [1.2 * 2.2 + 1e-9, 0, 0], [0, 0, 0], L)
    assert E == 0 and f[0] == 0


def test_lj_cut_indexed_rows():
    # IEpsISigCutAtom: eps and sigma looked up in the rows, NOT mixed
    eps = [[1.0, 0.3], [0.3, 2.0]]
    sig = [[1.0, 1.3], [1.3, 1.4]]   # arithmetic mean would be 1.2
    t = O.PairTable(O.LJ_CUT_II, eps, sig, [2.5, 2.0], [0, 1])
    rij = np.array([1.4, 0.7, -0.35])
    E, f = O.pair(3, t, 0, 1, rij, [0, 0, 0], [1e3] * 3)
    Ve, fe = _lj_cut_exact(rij, 0.3, 1.3, 2.5)
    _check(E, f, Ve, fe, 0.3, rij, 1.3)
    E, f = O.pair(3, t, 1, 1, rij, [0, 0, 0], [1e3] * 3)
    Ve, fe = _lj_cut_exact(rij, 2.0, 1.4, 2.0)
    _check(E, f, Ve, fe, 2.0, rij, 1.4)


@pytest.mark.parametrize("seed", range(3))
def test_lj_repulse_cut_at_sigma(seed):
    # EpsSigAtom: eps = sqrt(4 * 1) = 2, sigma = (1 + 1.4) / 2; unshifted,
    # zero from r = sigma on
    t = O.PairTable(O.LJ_REPULSE, [4.0, 1.0], [1.0, 1.4], None, [0, 1])
    rng = random.Random(seed)
    L = [1e3] * 2
    for _ in range(150):
        r = rng.uniform(0.8, 1.4)
        a = rng.uniform(0, 2 * math.pi)
        rij = np.array([r * math.cos(a), r * math.sin(a)])
        E, f = O.pair(2, t, 0, 1, rij, np.zeros(2), L)
        Ve, fe = _lj_cut_exact(rij, 2.0, 1.2, 1.0, shift=False)
        _check(E, f, Ve, fe, 2.0, rij, 1.2)
    E, f = O.pair(2, t, 0, 1, [1.2, 0], [0, 0], L)
    assert E == 0 and f[0] == 0
    E, f = O.pair(2, t, 0, 1, [1.1, 0], [0, 0], L)
    assert E > 0 and f[0] > 0


def test_repulsion_indexed_rows_harmonic():
    # IEpsISigExpAtom: eps, sigma from the rows; n = (2 + 2) / 2 = 2:
    # E = eps (1 - R/sig)^2 / 2, F = eps (1 - R/sig) / sig r_hat
    eps = [[1.0, 0.7], [0.7, 1.3]]
    sig = [[1.0, 1.15], [1.15, 1.4]]
    t = O.PairTable(O.REPULSION_II, eps, sig, [2.0, 2.0], [0, 1])
    for R in (Fr(1, 2), Fr(9, 10), Fr(11, 10)):
        rij = [R * 3 / 5, R * 4 / 5]
        E, f = O.pair(2, t, 0, 1, [float(c) for c in rij], [0, 0], [50.0, 50.0])
        x = 1 - R / Fr(115, 100)
        Ee = Fr(7, 10) * x * x / 2
        # 1 - R/sig loses digits as R nears sig: a relative error of a few
        # ulp of R/sig becomes ~ulp / x in E and F
        rel = 1e-15 / float(x)
        assert abs(E - float(Ee)) <= rel * float(Ee)
        fm = Fr(7, 10) * x / Fr(115, 100)
        assert abs(f[0] - float(fm * Fr(3, 5))) <= rel * float(fm)
        assert abs(f[1] - float(fm * Fr(4, 5))) <= rel * float(fm)
    # general exponent: n = (2.5 + 3.0) / 2
    t = O.PairTable(O.REPULSION_II, eps, sig, [2.5, 3.0], [0, 1])
    E, f = O.pair(2, t, 0, 1, [0.9, 0.0], [0, 0], [50.0, 50.0])
    n = mpmath.mpf(2.75)
    Ee = mpmath.mpf(0.7) * (1 - mpmath.mpf(0.9) / mpmath.mpf(1.15)) ** n / n
    assert abs(E - float(Ee)) <= 1e-14 * float(Ee)


# -------------------------------------------------------------- exclusions
def _numpy_pairs_excl(x, L, diam, skin, excl):
    x = np.asarray(x)
    out = set()
    for i in range(len(x)):
        d = x[i] - x[:i]
        d = d - L * np.round(d / L)
        r = np.sqrt(np.sum(d * d, axis=1))
        for j in np.nonzero(r < (diam[i] + diam[:i]) / 2 + skin)[0]:
            if (i, int(j)) not in excl:
                out.add((i, int(j)))
    return out


@pytest.mark.parametrize("ndim,n", [(3, 400), (2, 500)])
def test_nlist_exclusions_match_numpy(ndim, n):
    rng = np.random.default_rng(n)
    Lside = 9.0 if ndim == 3 else 22.0
    L = np.full(ndim, Lside)
    x = rng.uniform(-Lside, 2 * Lside, (n, ndim))
    diam = np.where(rng.random(n) < 0.5, 1.0, 1.4)
    full_f, full_l = O.nlist(x, L, diam, 0.3, brute=True)
    pick = rng.choice(len(full_f), 60, replace=False)
    ex = np.array([(full_l[p], full_f[p]) if p % 2 else (full_f[p], full_l[p])
                   for p in pick] + [(5, 5), (0, n - 1)], np.int32)
    excl = {(max(a, b), min(a, b)) for a, b in ex.tolist()}
    for brute in (True, False):
        f, l = O.nlist(x, L, diam, 0.3, brute=brute, ignore=ex)
        got = set(zip(f.tolist(), l.tolist()))
        assert got == _numpy_pairs_excl(x, L, diam, 0.3, excl)
        assert len(got) == len(full_f) - len(pick)


def test_ignore_forces_a_rebuild_and_is_counted_once():
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.5, 0.9, 0], [5, 5, 5]])
    t = O.PairTable(O.REPULSION, [1.0], [1.0], [2.0], [0, 0, 0, 0])
    s = O.System(np.ones(4), [20.0] * 3, t, [1.0] * 4, 0.3)
    s.set_state(x=x)
    assert s.update_list(False)          # the first update always rebuilds
    assert not s.update_list(False)      # nothing moved
    assert s.numpairs() == 3
    s.ignore(0, 1)
    s.ignore(1, 0)
    s.ignore(3, 3)
    assert s.ignore_size() == 2
    assert s.update_list(False)          # ignorechanged
    f, l = s.pairs()
    assert set(zip(f.tolist(), l.tolist())) == {(2, 0), (2, 1)}


# -------------------------------------------------------------------- NLCG
def _packing(n, ndim, phi, seed):
    rng = np.random.default_rng(seed)
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: n // 2]] = 1
    sig = np.array([1.0, 1.4])[types]
    Vs = np.sum(sig ** ndim) * (np.pi / 4 if ndim == 2 else np.pi / 6)
    Lside = (Vs / phi) ** (1 / ndim)
    x = rng.uniform(0, Lside, (n, ndim))
    t = O.PairTable(O.REPULSION, [1.0, 1.0], [1.0, 1.4], [2.0, 2.0], types)
    s = O.System(np.ones(n), [Lside] * ndim, t, sig, 0.3, ndim=ndim)
    s.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    return s


@pytest.mark.parametrize("ndim,P0", [(2, 1e-2), (3, 1e-2)])
def test_nlcg_reaches_mechanical_equilibrium_at_goal_pressure(ndim, P0):
    """packer.cpp's use (src/bin/packer.cpp:102-147): minimise H = U + P0 V
    over positions and box; at the minimum the virial pressure is P0 and
    every force vanishes."""
    s = _packing(48, ndim, 0.5, seed=ndim)
    c = O.NLCG(s, 0.1, P0)
    c.set_limits(stepmax=1e-2)
    H0 = c.observables()["H"]
    for _ in range(40):
        c.timestep(100)
        x, _, _, f = s.state()
        if np.abs(f).max() < 1e-11 and abs(c.observables()["P"] / P0 - 1) < 1e-9:
            break
    obs = c.observables()
    _, _, _, f = s.state()
    assert abs(obs["P"] / P0 - 1) < 1e-9, obs
    assert np.abs(f).max() < 1e-11
    assert obs["H"] < H0
    # the state is a packing under load: overlaps remain
    U = s.potential_energy()
    assert U > 0


def test_nlcg_fixed_L_changes_the_first_edge_only():
    s = _packing(40, 2, 0.5, seed=7)
    L0 = s.box()
    c = O.NLCG(s, 1e-2, 1e-2, fixed_L=True)
    c.timestep(300)
    L1 = s.box()
    assert L1[1] == L0[1] and L1[0] != L0[0]


def test_nlcg_reset_and_descend():
    """reset() sets v = a = f and vl = al = fl (the steepest-descent
    direction); descend() then steps along it by dt: the box edge grows by
    exp(dt fl / (kappa NDIM)) exactly and every atom moves by Lfac x + dt f."""
    s = _packing(30, 3, 0.3, seed=3)
    c = O.NLCG(s, 1e-3, 1.0)
    c.reset()
    x0, v0, a0, f0 = s.state()
    info = c.info()
    assert np.array_equal(v0, f0) and np.array_equal(a0, f0)
    assert info["vl"] == info["al"] == info["fl"]
    L0 = s.box()
    c.descend()
    fl = c.info()["fl"]
    _, _, _, f1 = s.state()
    x1 = s.state()[0]
    Lfac = math.exp(1e-3 * fl / (10.0 * 3))
    assert np.allclose(s.box(), L0 * Lfac, rtol=1e-15, atol=0)
    assert np.allclose(x1, x0 * Lfac + 1e-3 * f1, rtol=1e-14, atol=1e-14)


def _anchored_chain(seed, kmax):
    """Four atoms on a line between two fixed anchors, every pair inside its
    harmonic-repulsion contact: while the order along the line holds, the
    energy is an exact quadratic in the four mobile x coordinates (positive
    definite through the strong anchors).  The box is frozen: kappa = 1e30
    makes Lfac = exp(dx vl / (kappa NDIM)) exactly 1."""
    n = 6
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 3))
    x[:, 0] = [0.0, 1.2, 2.4, 3.6, 4.8, 6.0]
    x[1:5, 0] += rng.uniform(-0.1, 0.1, 4)
    sig = np.array([7.0, 7.5, 8.0, 8.5, 9.0, 7.2])
    eps = np.array([50.0, 2.0, 0.5, 1.5, 3.0, 50.0])
    m = np.array([np.inf, 1.0, 1.0, 1.0, 1.0, np.inf])
    table = O.PairTable(O.REPULSION, eps, sig, np.full(n, 2.0), np.arange(n, dtype=np.int32))
    s = O.System(m, np.full(3, 100.0), table, sig, 0.5, ndim=3)
    s.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    s.initialize()
    c = O.NLCG(s, 1e-3, 0.0, kappa=1e30, kmax=kmax)
    c.set_limits(alphamax=0.0, afrac=0.0, dxmax=0.0, stepmax=0.0, maxdV=0.0)
    c.reset()
    return s, c


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_nlcg_conjugate_gradient_terminates_on_a_quadratic(seed):
    """On an exact quadratic in d unknowns, conjugate gradients with exact
    line searches reach the minimum in at most d steps; the secant search
    is exact on a quadratic (its derivative along the line is linear).  So
    CollectionNLCG::timestep (src/collection.cpp:685-835) must bring the
    forces to rounding level in 4 steps here, while the same steps with the
    direction reset every step (kmax = 1: steepest descent) must not: this
    pins the secant step, the Polak-Ribiere beta's denominator and its
    f_new.f_new term, and v = a + beta v.  (On a quadratic with exact line
    searches f_new.f_old = 0, so the beta numerator's -f_new.f_old term is
    not pinned by this test.)"""
    def fmax(sys):  # the mobile atoms' forces (the anchors' are not minimised)
        return np.abs(sys.state()[3][1:5]).max()

    s, c = _anchored_chain(seed, kmax=1000)
    f0 = fmax(s)
    c.timestep(3)
    assert fmax(s) > 1e-6 * f0
    c.timestep(1)
    assert fmax(s) < 1e-13 * f0
    assert np.all(np.diff(s.state()[0][:, 0]) > 0)  # the order held: the quadratic was exact
    s2, c2 = _anchored_chain(seed, kmax=1)
    c2.timestep(4)
    assert fmax(s2) > 1e-6 * f0

@pytest.mark.parametrize("seed", [0, 1, 3])
def test_nlcg_beta_is_polak_ribiere(seed):
    """The second CollectionNLCG step's beta is Polak and Ribiere's
    coefficient (the reference names the method at src/collection.cpp:696
    and forms it at :802-815): beta = f2.(f2 - f1) / f1.f1, with f1 the forces
    after the first step and f2 after the second, in the step's inner
    product (sum over mobile atoms; the box terms vanish with kappa = 1e30,
    which also freezes the box, so the length-squared normaliser cancels).
    After the first step f2.f1 is not zero here (harmonic repulsion is not
    quadratic in the positions), so Fletcher and Reeves' f2.f2 / f1.f1, the
    same formula without the -f2.f1 term, differs by over 20 %.  The clamp to
    [0, 1] gives betaused."""
    s = _packing(30, 3, 0.6, seed=seed)
    c = O.NLCG(s, 1e-3, 0.0, kappa=1e30)
    c.set_limits(alphamax=0.0, afrac=0.0, dxmax=0.0, stepmax=0.0, maxdV=0.0)
    c.reset()
    c.timestep(1)
    f1 = s.state()[2].copy()  # a = f after a step
    c.timestep(1)
    f2 = s.state()[2].copy()
    info = c.info()
    beta_pr = (np.sum(f2 * f2) - np.sum(f2 * f1)) / np.sum(f1 * f1)
    beta_fr = np.sum(f2 * f2) / np.sum(f1 * f1)
    assert abs(beta_fr - beta_pr) > 0.2 * abs(beta_pr)
    assert info["beta"] == pytest.approx(beta_pr, rel=1e-12)
    assert info["betaused"] == min(max(beta_pr, 0.0), 1.0) or \
        info["betaused"] == pytest.approx(min(max(beta_pr, 0.0), 1.0), rel=1e-12)
