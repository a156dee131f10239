"""The benchmark configurations at their full sizes, on the device, against
the oracle where a comparison is feasible in a test (the oracle is one CPU
core) and through size-independent properties where it is not:

  * 1,048,576-atom LJ, FP64 (plain): pair set, forces, U and W at step 0 and
    after 10 Verlet steps against the oracle's cell-list restatement;
  * the same system in mixed precision: per-atom force error against the
    oracle within a bound of FP32 arithmetic (2e-5 of sum_j |f_ij| plus 1e-6
    of the scale of the force's two terms, which cancel near the LJ minimum)
    and of the FP32-rounded displacement (the force's sensitivity to it times
    the largest rounding of a staged offset); both precisions conserve
    energy over
    10^4 NVE steps (relative drift < 1e-4 after a 1000-step equilibration);
  * 262,144-disk 2D bidisperse relaxation: pair set and forces at step 0
    against the oracle, then damped dynamics to mean |F| < 1e-10;
  * 4096 batched packings (8.5 M atoms): 2000 damped steps in mixed
    precision with property checks, and in FP64 the first 10 steps of a
    sample of systems against their own oracle systems.
"""
import numpy as np
import pytest

from oracle import oracle as O
from parm_amd import capi, workloads
from tests import systems as S

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N_LJ = 1 << 20


def _device_or_fail():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible: the -m gpu tests need an MI355X")


@pytest.fixture(scope="module")
def lj_liquid():
    _device_or_fail()
    print("building the 1M-atom LJ liquid", flush=True)
    return workloads.lj_liquid(N_LJ, rho=0.8, T=1.0, min_sep=0.8, seed=2024)


def _lj_device(x, v, L, prec):
    n = len(x)
    s = capi.System(3, np.ones(n), prec)
    s.set_box(L)
    s.nlist_set(0.3, np.full(n, 2.5))
    s.set_lj([[1.0]], [1.0], [2.5], np.zeros(n, np.int32))
    s.set_state(x=x, v=v, a=np.zeros_like(x))
    s.initialize()
    return s


def _lj_table(n):
    return O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], np.zeros(n, np.int32))


def _lj_mixed_scales(x, L, first, last, cut, chunk=1 << 22):
    """Per atom, over its pairs within the cut (eps = sigma = 1, s = r^-6):
    the sum of 12 s (s + 1) / r, the scale of the force's two terms (FP32
    rounds each; they cancel near the minimum), and the sum of |f'(r)| +
    |f(r)| / r, which bounds how far the pair force vector moves per unit
    change of the displacement."""
    term = np.zeros(len(x))
    sens = np.zeros(len(x))
    for a in range(0, len(first), chunk):
        f, l = first[a:a + chunk], last[a:a + chunk]
        d = x[f] - x[l]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        s6 = r ** -6.0
        inside = r < cut
        t = np.where(inside, 12.0 * s6 * (s6 + 1.0) / r, 0.0)
        fmag = 24.0 * s6 * (2.0 * s6 - 1.0) / r
        dfdr = 24.0 * (26.0 * s6 * s6 - 7.0 * s6) / (r * r)
        g = np.where(inside, np.abs(dfdr) + np.abs(fmag) / r, 0.0)
        for out, v in ((term, t), (sens, g)):
            out += np.bincount(f, v, len(x)) + np.bincount(l, v, len(x))
    return term, sens


def test_config2_lj_1M_fp64_parity_over_10_steps(lj_liquid):
    x, v, L = lj_liquid
    n = len(x)
    print("config 2: device system", flush=True)
    s = _lj_device(x, v, L, capi.FP64)
    print("config 2: oracle system (cell-list build)", flush=True)
    ref = O.System(np.ones(n), L, _lj_table(n), np.full(n, 2.5), 0.3, brute=False)
    ref.set_state(x=x, v=v, a=np.zeros_like(x))
    ref.initialize()
    for step in (0, 10):
        if step:
            print("config 2: 10 Verlet steps", flush=True)
            s.verlet(0.002, 10)
            ref.verlet(0.002, 10)
        assert s.which() == ref.which()
        f_dev, l_dev = s.pairs()
        first, last = ref.pairs()
        assert np.array_equal(f_dev, first) and np.array_equal(l_dev, last), step
        xr, vr, _, fr = ref.state()
        xd, vd, _, fd = s.state()
        assert np.abs(xd - xr).max() <= 1e-12 * np.abs(xr).max(), step
        assert np.abs(vd - vr).max() <= 1e-11 * np.abs(vr).max(), step
        chk = O.forces(xr, L, _lj_table(n), first, last)
        err = S.normalised_force_error(fd, chk["f"], chk["fabs"])
        if step:
            assert err <= 1e-10, (step, err)
        else:
            S.assert_fp64_forces(fd, chk, xr, L, _lj_table(n), first, last)
        sums = s.pair_sums()
        assert abs(sums["U"] - chk["U"]) <= 1e-12 * abs(chk["U"]), step
        assert abs(sums["W"] - chk["W"]) <= 1e-12 * abs(chk["W"]), step
        print(f"config 2 step {step}: {len(first)} pairs, force error {err:.2e}", flush=True)
    s.close()


def test_config3_mixed_force_error_and_drift(lj_liquid):
    x, v, L = lj_liquid
    n = len(x)
    sm = _lj_device(x, v, L, capi.MIXED)
    first, last = O.nlist(x, L, np.full(n, 2.5), 0.3)
    chk = O.forces(x, L, _lj_table(n), first, last)
    d = np.sqrt(((sm.forces() - chk["f"]) ** 2).sum(1))
    term, sens = _lj_mixed_scales(x, L, first, last, 2.5)
    # FP32 pair arithmetic (2e-5 of sum_j |f_ij| plus 1e-6 of the terms'
    # scale), and the FP32 displacement: the list stages each atom's offset
    # from its tile's origin in FP32, a minimum image (|o| <= L/2 per
    # coordinate) rounded by at most 2^-24 |o|; a pair displacement differences
    # two of them, so it moves by at most delta = 2 sqrt(3) 2^-24 L/2.
    delta = 2.0 * np.sqrt(3.0) * 2.0 ** -24 * (L.max() / 2)
    bar = 2e-5 * chk["fabs"] + 1e-6 * term + delta * sens
    rel = d / chk["fabs"]
    print(f"config 3: mixed per-atom force error max {d.max():.3e} rms "
          f"{np.sqrt(np.mean(d * d)):.3e}; normalised by sum_j |f_ij| max {rel.max():.3e} "
          f"rms {np.sqrt(np.mean(rel * rel)):.3e}; max error / bound "
          f"{(d / bar).max():.3f}", flush=True)
    assert np.all(d <= bar), float((d / bar).max())
    sm.close()
    for prec in (capi.FP64, capi.MIXED):
        print(f"config 3: 10^4-step NVE drift, precision {prec}", flush=True)
        s = _lj_device(x, v, L, prec)
        s.verlet(0.002, 1000)
        E = [s.pair_sums()["U"] + s.kinetic_energy()]
        for _ in range(100):
            s.verlet(0.002, 100)
            E.append(s.pair_sums()["U"] + s.kinetic_energy())
        E = np.asarray(E)
        drift = np.max(np.abs(E - E[0])) / abs(E[0])
        print(f"config 3: precision {prec} drift {drift:.2e}", flush=True)
        assert drift < 1e-4, drift
        s.close()


def test_config4_relaxation_262144_disks():
    _device_or_fail()
    n = 262144
    x, types, L = workloads.bidisperse_packing(n, 0.84, 2026, 2)
    sig = np.array([1.0, 1.4])
    table = O.PairTable(O.REPULSION, [1.0, 1.0], sig, [2.0, 2.0], types)
    # step 0 in FP64 against the oracle
    s = capi.System(2, np.ones(n), capi.FP64)
    s.set_box(L)
    s.nlist_set(0.4, sig[types])
    s.set_repulsion([1.0, 1.0], sig, [2.0, 2.0], types)
    s.set_state(x=x)
    s.update_list(True)
    s.set_forces(False)
    first, last = O.nlist(x, L, sig[types], 0.4)
    f_dev, l_dev = s.pairs()
    assert np.array_equal(f_dev, first) and np.array_equal(l_dev, last)
    chk = O.forces(x, L, table, first, last)
    assert S.normalised_force_error(s.forces(), chk["f"], chk["fabs"]) <= 1e-12
    s.close()
    # the relaxation, mixed precision, to mean |F| < 1e-10
    s = capi.System(2, np.ones(n), capi.MIXED)
    s.set_box(L)
    s.nlist_set(0.4, sig[types])
    s.set_repulsion([1.0, 1.0], sig, [2.0, 2.0], types)
    s.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    s.initialize()
    steps = 0
    fmean = s.force_stats()[1]
    while fmean >= 1e-10 and steps < 20000:
        s.damped(0.3, 0.02, 250)
        steps += 250
        fmean = s.force_stats()[1]
    print(f"config 4: mean |F| {fmean:.2e} after {steps} steps", flush=True)
    assert fmean < 1e-10
    assert np.all(np.isfinite(s.positions()))
    s.close()


def _packings(nsys, seed):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(64, 4096 + 1, nsys)
    types, xs, Ls = [], [], []
    for n in sizes:
        t = np.zeros(n, np.int32)
        t[rng.permutation(n)[: n // 2]] = 1
        sig = np.array([1.0, 1.4])[t]
        Ls.append(float((np.sum(np.pi / 6 * sig ** 3) / 0.64) ** (1 / 3)))
        types.append(t)
        xs.append(rng.uniform(0, Ls[-1], (n, 3)))
    return sizes, types, xs, np.asarray(Ls)


def test_config5_batched_packings_4096():
    _device_or_fail()
    sizes, types, xs, Ls = _packings(4096, 2025)
    t_all = np.concatenate(types)
    x_all = np.concatenate(xs)
    sig = np.array([1.0, 1.4])
    ntot = int(sizes.sum())

    def batch(prec):
        b = capi.Batch(3, sizes, np.repeat(Ls[:, None], 3, 1), 0.2, prec)
        b.set_repulsion([1.0, 1.0], sig, [2.0, 2.0], t_all)
        b.set_state(x=x_all, v=np.zeros_like(x_all), a=np.zeros_like(x_all))
        b.initialize()
        return b

    print(f"config 5: {len(sizes)} systems, {ntot} atoms, 2000 damped steps", flush=True)
    b = batch(capi.MIXED)
    U0 = b.pair_sums()["U"]
    w0, _ = b.lists()
    b.damped(0.01, 1.0, 2000)
    sums = b.pair_sums()
    w1, np1 = b.lists()
    x1 = b.state()[0]
    b.close()
    assert np.all(np.isfinite(x1))
    assert np.all(sums["U"] >= 0) and sums["U"].sum() < 0.05 * U0.sum()
    assert np.all(w1 > w0)                 # every system rebuilt along the way
    assert np.all(sums["overlaps"] <= sums["contacts"])
    assert np.all(np1 >= 0)
    # FP64: a sample of systems (smallest, largest, some in between) follow
    # their own oracle systems over the first 10 steps
    b = batch(capi.FP64)
    b.damped(0.01, 1.0, 10)
    xb = b.state()[0]
    b.close()
    off = np.concatenate([[0], np.cumsum(sizes)])
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 0, 1, 2047, 4095})
    for q in sample:
        n = int(sizes[q])
        sl = slice(off[q], off[q + 1])
        tab = O.PairTable(O.REPULSION, [1.0, 1.0], sig, [2.0, 2.0], types[q])
        ref = O.System(np.ones(n), [Ls[q]] * 3, tab, sig[types[q]], 0.2)
        ref.set_state(x=xs[q], v=np.zeros((n, 3)), a=np.zeros((n, 3)))
        ref.initialize()
        ref.damped(0.01, 1.0, 10)
        xr = ref.state()[0]
        assert np.abs(xb[sl] - xr).max() <= 1e-12 * np.abs(xr).max(), q
