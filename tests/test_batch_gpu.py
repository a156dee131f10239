"""Batched systems (parm_batch_*) vs the CPU oracle, one oracle System per
system, through the C ABI.

Every system of a batch is what the reference builds with one AtomVec,
OriginBox, NeighborList, NListed and Collection, so each is checked against
its own oracle System on the same inputs:
  * pair sets: bit-identical per system, also after device-side rebuilds;
  * which(): the per-system rebuild counts equal the oracle's;
  * FP64 trajectories: 1e-12 relative over the first 10 steps (only the force
    summation order differs), damped and Verlet;
  * graph replay == eager launches, bit for bit, over several calls;
  * mixed precision: forces within 2e-5 of the FP64 batch (normalised per
    atom as in the single-system tests).
The batch holds ragged sizes including 1- and 2-atom systems and boxes too
small for a 3x3x3 cell grid.
"""
import numpy as np
import pytest

from oracle import oracle as O
from parm_amd import capi
from tests import systems as S

pytestmark = pytest.mark.gpu


def _device_or_fail():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible: the -m gpu tests need an MI355X")


def packings(sizes, phi, ndim, seed):
    """Per system: bidisperse (1.0 / 1.4, 50/50) harmonic spheres at packing
    fraction phi, uniform random unwrapped positions."""
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        types = np.zeros(n, np.int32)
        types[rng.permutation(n)[: n // 2]] = 1
        sig = np.array([1.0, 1.4])[types]
        vol = np.sum(np.pi / 6 * sig ** 3) if ndim == 3 else np.sum(np.pi / 4 * sig ** 2)
        L = float((vol / phi) ** (1.0 / ndim))
        x = rng.uniform(0, L, (n, ndim)) + rng.integers(-2, 3, (n, ndim)) * L
        v = rng.normal(size=(n, ndim)) * 0.3
        out.append(dict(n=n, L=L, x=x, v=v, types=types, sig=sig))
    return out


REP = dict(eps=[1.0, 1.0], sigma=[1.0, 1.4], third=[2.0, 2.0])


def make_batch(systems, ndim, skin, prec=capi.FP64, lj=False):
    b = capi.Batch(ndim, [s["n"] for s in systems],
                   np.array([[s["L"]] * ndim for s in systems]), skin, prec)
    types = np.concatenate([s["types"] for s in systems])
    if lj:
        b.set_lj([[1.0, 1.0], [1.0, 1.0]], [1.0, 1.4], [2.5, 2.5], types)
    else:
        b.set_repulsion(REP["eps"], REP["sigma"], REP["third"], types)
    b.set_state(x=np.concatenate([s["x"] for s in systems]),
                v=np.concatenate([s["v"] for s in systems]))
    return b


def oracle_systems(systems, ndim, skin, lj=False):
    refs = []
    for s in systems:
        if lj:
            t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0, 1.0], [1.0, 1.0]], [1.0, 1.4],
                            [2.5, 2.5], s["types"])
            diam = np.array([2.5, 3.5])[s["types"]]
        else:
            t = O.PairTable(O.REPULSION, REP["eps"], REP["sigma"], REP["third"], s["types"])
            diam = s["sig"]
        r = O.System(np.ones(s["n"]), [s["L"]] * ndim, t, diam, skin, ndim=ndim,
                     brute=True)
        r.set_state(x=s["x"], v=s["v"], a=np.zeros_like(s["x"]))
        refs.append(r)
    return refs


SIZES = [1, 2, 5, 64, 37, 300, 129, 700, 64, 3, 1000, 250]


def _compare(b, refs, systems, xtol, tag):
    xd, vd, ad, fd = b.state()
    w, npairs = b.lists()
    off = b.offsets
    for s, r in enumerate(refs):
        sl = slice(off[s], off[s + 1])
        xr, vr, ar, fr = r.state()
        assert w[s] == r.which(), (tag, s)
        assert npairs[s] == r.numpairs(), (tag, s)
        scale = max(1.0, np.abs(xr).max())
        assert np.abs(xd[sl] - xr).max() <= xtol * scale, (tag, s)
        assert np.abs(vd[sl] - vr).max() <= 10 * xtol * max(1.0, np.abs(vr).max()), (tag, s)
        assert np.abs(fd[sl] - fr).max() <= 100 * xtol * max(1.0, np.abs(fr).max()), (tag, s)


@pytest.mark.parametrize("ndim", [3, 2])
def test_batch_pair_sets_bit_identical(ndim):
    _device_or_fail()
    print(f"batch pair sets ndim={ndim}", flush=True)
    systems = packings(SIZES, 0.64 if ndim == 3 else 0.84, ndim, seed=5 + ndim)
    b = make_batch(systems, ndim, 0.2)
    refs = oracle_systems(systems, ndim, 0.2)
    b.initialize()
    for r in refs:
        r.initialize()
    w, npairs = b.lists()
    for s, r in enumerate(refs):
        f_dev, l_dev = b.pairs(s)
        f_ref, l_ref = r.pairs()
        assert np.array_equal(f_dev, f_ref) and np.array_equal(l_dev, l_ref), s
        assert npairs[s] == len(f_ref) and w[s] == r.which() == 1, s
    U, c = b.sums()
    for s, r in enumerate(refs):
        assert abs(U[s] - r.potential_energy()) <= 1e-12 * max(1.0, abs(r.potential_energy()))


@pytest.mark.parametrize("integrator,ndim", [("damped", 3), ("verlet", 3), ("damped", 2)])
def test_batch_trajectory_matches_oracle(integrator, ndim):
    _device_or_fail()
    print(f"batch trajectory {integrator} ndim={ndim}", flush=True)
    lj = integrator == "verlet"
    if lj:
        systems = []
        for k, n in enumerate([1, 2, 40, 300, 125, 512]):
            if n <= 2:
                # a lattice of 1-2 sites at density 0.8 puts atoms 0.55 sigma
                # apart, where dt = 0.004 is unstable; place them by hand
                L = np.full(3, 3.0)
                x = np.array([[0.2, 0.3, 0.1], [1.3, 0.5, -0.2]])[:n] + 7.0 * k
            else:
                x, L = S.jittered_lattice(n, 0.8, 3, seed=40 + k)
            types = np.zeros(n, np.int32)
            systems.append(dict(n=n, L=float(L[0]), x=x,
                                v=S.gaussian_velocities(n, 3, 1.0, seed=k),
                                types=types, sig=np.ones(n)))
        skin = 0.3
    else:
        systems = packings(SIZES, 0.64 if ndim == 3 else 0.84, ndim, seed=17 + ndim)
        skin = 0.2
    b = make_batch(systems, ndim, skin, lj=lj)
    refs = oracle_systems(systems, ndim, skin, lj=lj)
    b.initialize()
    for r in refs:
        r.initialize()
    dt = 0.004 if lj else 0.01
    for k in (1, 2, 3, 4):
        if lj:
            b.verlet(dt, k)
        else:
            b.damped(dt, 1.0, k)
        for r in refs:
            if lj:
                r.verlet(dt, k)
            else:
                r.damped(dt, 1.0, k)
        _compare(b, refs, systems, 1e-12, k)
    # a longer graph-replayed run: device-side rebuilds in every system
    n2 = 41
    if lj:
        b.verlet(dt, n2)
    else:
        b.damped(dt, 1.0, n2)
    for r in refs:
        if lj:
            r.verlet(dt, n2)
        else:
            r.damped(dt, 1.0, n2)
    _compare(b, refs, systems, 1e-9, "long")
    w, _ = b.lists()
    assert w.max() > 1
    for s, r in enumerate(refs):
        f_dev, l_dev = b.pairs(s)
        f_ref, l_ref = r.pairs()
        assert np.array_equal(f_dev, f_ref) and np.array_equal(l_dev, l_ref), s


def test_batch_list_update_order_with_zero_skin():
    """As the single-system test: with skin 0 every step rebuilds, and the
    forces at x_{q+1} must use the list tested at x_q (the reference's
    timestep order), in every system of the batch and through graph replay."""
    _device_or_fail()
    print("batch zero skin", flush=True)
    systems = []
    for k, n in enumerate([300, 125, 512, 64]):
        x, L = S.jittered_lattice(n, 0.8, 3, seed=60 + k)
        systems.append(dict(n=n, L=float(L[0]), x=x,
                            v=S.gaussian_velocities(n, 3, 1.0, seed=70 + k),
                            types=np.zeros(n, np.int32), sig=np.ones(n)))
    b = make_batch(systems, 3, 0.0, lj=True)
    refs = oracle_systems(systems, 3, 0.0, lj=True)
    b.initialize()
    for r in refs:
        r.initialize()
    for k in (1, 4, 21):
        b.verlet(0.005, k)
        for r in refs:
            r.verlet(0.005, k)
        _compare(b, refs, systems, 1e-11, k)
    w, _ = b.lists()
    assert np.all(w == 1 + 26)


def test_batch_graph_replay_equals_eager(monkeypatch):
    _device_or_fail()
    print("batch graph vs eager", flush=True)
    # one lanes-per-atom layout for both (the autotuner could pick two)
    monkeypatch.setenv("PARM_TPA", "4")
    systems = packings([400, 90, 1, 1500, 33, 800], 0.64, 3, seed=99)
    bg = make_batch(systems, 3, 0.2)
    be = make_batch(systems, 3, 0.2)
    bg.initialize()
    be.initialize()
    for k in (7, 20, 1, 33):
        bg.damped(0.01, 1.0, k, graph=True)
        be.damped(0.01, 1.0, k, graph=False)
        for a, e in zip(bg.state(), be.state()):
            assert np.array_equal(a, e), k
        assert np.array_equal(bg.lists()[0], be.lists()[0])
    # changing the step parameters re-captures
    bg.damped(0.005, 0.5, 9, graph=True)
    be.damped(0.005, 0.5, 9, graph=False)
    for a, e in zip(bg.state(), be.state()):
        assert np.array_equal(a, e)


def test_batch_mixed_precision_forces():
    _device_or_fail()
    print("batch mixed precision", flush=True)
    systems = packings([500, 64, 2000, 129], 0.64, 3, seed=3)
    b64 = make_batch(systems, 3, 0.2)
    b32 = make_batch(systems, 3, 0.2, prec=capi.MIXED)
    b64.initialize()
    b32.initialize()
    f64 = b64.state()[3]
    f32 = b32.state()[3]
    # per-atom normaliser: sum_j |f_ij| from the oracle's pair forces
    fabs = np.zeros(len(f64))
    off = b64.offsets
    for s, sy in enumerate(systems):
        t = O.PairTable(O.REPULSION, REP["eps"], REP["sigma"], REP["third"], sy["types"])
        first, last = O.nlist(sy["x"], [sy["L"]] * 3, sy["sig"], 0.2, brute=True)
        fabs[off[s]:off[s + 1]] = O.forces(sy["x"], [sy["L"]] * 3, t, first, last)["fabs"]
    err = S.normalised_force_error(f32, f64, fabs)
    assert err < 2e-5, err
    b64.damped(0.01, 1.0, 50)
    b32.damped(0.01, 1.0, 50)
    U64, _ = b64.sums()
    U32, _ = b32.sums()
    assert np.all(np.abs(U32 - U64) <= 1e-3 * np.maximum(U64, 1e-3))


def test_batch_errors():
    _device_or_fail()
    with pytest.raises(ValueError):
        capi.Batch(3, [10, 0], [[3.0] * 3] * 2, 0.2)
    b = capi.Batch(3, [10, 4], [[3.0] * 3, [2.0] * 3], 0.2)
    with pytest.raises(RuntimeError, match="no pair interaction"):
        b.initialize()
    with pytest.raises(ValueError, match="type out of range"):
        b.set_repulsion([1.0], [1.0], [2.0], np.full(14, 3, np.int32))
