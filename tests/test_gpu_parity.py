"""Device path vs the CPU oracle, through the C ABI (libparm_hip.so).

Bars:
  * neighbour-pair set: bit-identical (same pairs, same reference order);
  * FP64 forces: |F_dev - F_ref| <= 1e-12 * sum_j |f_ij| per atom;
  * FP64 energy / virial / stress: 1e-12 relative; contacts / overlaps exact;
  * integrators: the device step reproduces the oracle's trajectory to
    1e-12 relative over 10 steps (only force summation order differs);
  * mixed precision: forces within 2e-5 of FP64 (normalised as above).
"""
import numpy as np
import pytest

from oracle import oracle as O
from parm_amd import capi
from tests import systems as S

pytestmark = pytest.mark.gpu


def _device_or_skip():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible: the -m gpu tests need an MI355X")


def _table(kind, p):
    return O.PairTable(kind, p["eps"], p["sigma"], p["third"], p["types"])


def _make(kind, x, L, p, skin, prec=capi.FP64, masses=None, members=None):
    n, ndim = x.shape
    s = capi.System(ndim, np.ones(n) if masses is None else masses, prec)
    s.set_box(L)
    if members is None:
        s.nlist_set(skin, p["diam"])
    else:
        s.nlist_set(skin, p["diam"][members], members)
    if kind == O.LJ_ATTRACT_REPULSE:
        s.set_lj(p["eps"], p["sigma"], p["third"], p["types"])
    else:
        s.set_repulsion(p["eps"], p["sigma"], p["third"], p["types"])
    s.set_state(x=x)
    return s


CASES = [
    # kind, ndim, n, rho, skin, bidisperse
    (O.LJ_ATTRACT_REPULSE, 3, 1000, 0.8, 0.4, False),
    (O.LJ_ATTRACT_REPULSE, 3, 4096, 0.8, 0.3, False),
    (O.REPULSION, 3, 3000, 1.0, 0.2, True),
    (O.REPULSION, 2, 5000, 1.0, 0.2, True),
    (O.LJ_ATTRACT_REPULSE, 2, 2000, 0.6, 0.3, False),
]


@pytest.mark.parametrize("kind,ndim,n,rho,skin,bi", CASES)
def test_pair_set_bit_identical(kind, ndim, n, rho, skin, bi):
    _device_or_skip()
    print(f"pair-set case kind={kind} ndim={ndim} n={n}", flush=True)
    x, L = S.jittered_lattice(n, rho, ndim, seed=n + ndim)
    p = S.bidisperse_repulsion(n, n) if bi else S.lj_params(n)
    s = _make(kind, x, L, p, skin)
    assert s.update_list(True)
    f_dev, l_dev = s.pairs()
    f_ref, l_ref = O.nlist(x, L, p["diam"], skin, brute=(n <= 3000))
    assert s.numpairs() == len(f_ref)
    assert np.array_equal(f_dev, f_ref) and np.array_equal(l_dev, l_ref)
    # the device keeps the caller's positions bit for bit through its sort
    assert np.array_equal(s.positions(), x)


def test_pair_set_subset_members_and_tiny_box():
    _device_or_skip()
    rng = np.random.default_rng(4)
    # box smaller than 3 list radii (cells collapse), components near L/2
    n = 300
    L = np.array([4.0, 3.7, 4.3])
    x = rng.uniform(-10, 10, (n, 3))
    x[1] = x[0] + np.array([2.0, 0.1, 0.0])      # |dx| = L/2 exactly (tie)
    x[2] = x[0] + np.array([6.0, 0.0, 0.2])      # L + L/2
    members = rng.permutation(n)[:250].astype(np.int32)
    members = np.concatenate([[0, 1, 2], members[~np.isin(members, [0, 1, 2])]]).astype(np.int32)
    p = S.bidisperse_repulsion(n, 3)
    s = _make(O.REPULSION, x, L, p, 0.45, members=members)
    s.update_list(True)
    f_dev, l_dev = s.pairs()
    f_ref, l_ref = O.nlist(x, L, p["diam"][members], 0.45, member=members, brute=True)
    assert np.array_equal(f_dev, f_ref) and np.array_equal(l_dev, l_ref)


@pytest.mark.parametrize("kind,ndim,n,rho,skin,bi", CASES)
def test_fp64_forces_energy_virial(kind, ndim, n, rho, skin, bi):
    _device_or_skip()
    print(f"forces case kind={kind} ndim={ndim} n={n}", flush=True)
    x, L = S.jittered_lattice(n, rho, ndim, seed=7 * n + ndim, jitter=0.12)
    p = S.bidisperse_repulsion(n, n) if bi else S.lj_params(n)
    s = _make(kind, x, L, p, skin)
    s.update_list(True)
    s.set_forces(False)
    f_dev = s.forces()
    first, last = O.nlist(x, L, p["diam"], skin)
    ref = O.forces(x, L, _table(kind, p), first, last)
    S.assert_fp64_forces(f_dev, ref, x, L, _table(kind, p), first, last)
    sums = s.pair_sums()
    assert abs(sums["U"] - ref["U"]) <= 1e-12 * max(abs(ref["U"]), 1e-300) + 1e-13
    assert abs(sums["W"] - ref["W"]) <= 1e-12 * max(abs(ref["W"]), 1e-300) + 1e-13
    scale = np.abs(ref["S"]).max() + 1e-300
    assert np.abs(sums["S"] - ref["S"]).max() <= 1e-12 * scale
    assert sums["contacts"] == ref["contacts"]
    assert sums["overlaps"] == ref["overlaps"]


def test_repulsion_general_exponent_and_multi_type_lj():
    _device_or_skip()
    n = 2000
    x, L = S.jittered_lattice(n, 0.9, 3, seed=5, jitter=0.15)
    p = S.bidisperse_repulsion(n, 5, eps=(1.0, 2.0), exponent=(2.5, 3.0))
    s = _make(O.REPULSION, x, L, p, 0.3)
    s.update_list(True)
    s.set_forces(True)
    first, last = O.nlist(x, L, p["diam"], 0.3)
    ref = O.forces(x, L, _table(O.REPULSION, p), first, last)
    S.assert_fp64_forces(s.forces(), ref, x, L, _table(O.REPULSION, p), first, last)
    # two-type LJ with an eps <= 0 (purely repulsive) type pair
    types = (np.arange(n) % 2).astype(np.int32)
    q = dict(eps=[[1.0, 0.7], [0.7, -1.5]], sigma=[1.0, 1.2], third=[2.5, 2.2],
             types=types, diam=np.where(types == 1, 1.2 * 2.2, 2.5))
    s2 = _make(O.LJ_ATTRACT_REPULSE, x, L, q, 0.3)
    s2.update_list(True)
    s2.set_forces(False)
    first, last = O.nlist(x, L, q["diam"], 0.3)
    ref = O.forces(x, L, _table(O.LJ_ATTRACT_REPULSE, q), first, last)
    S.assert_fp64_forces(s2.forces(), ref, x, L, _table(O.LJ_ATTRACT_REPULSE, q), first, last)
    assert abs(s2.pair_sums()["U"] - ref["U"]) <= 1e-12 * abs(ref["U"])


def _oracle_system(kind, x, L, p, skin, masses):
    n, ndim = x.shape
    return O.System(masses, L, _table(kind, p), p["diam"], skin, ndim=ndim)


@pytest.mark.parametrize("kind,ndim,n,integrator", [
    (O.LJ_ATTRACT_REPULSE, 3, 1000, "verlet"),
    (O.REPULSION, 2, 4000, "damped"),
    (O.REPULSION, 3, 2000, "verlet"),
])
def test_trajectory_matches_oracle(kind, ndim, n, integrator):
    _device_or_skip()
    print(f"trajectory case kind={kind} ndim={ndim} n={n} {integrator}", flush=True)
    rho = 0.8 if kind == O.LJ_ATTRACT_REPULSE else 1.0
    x, L = S.jittered_lattice(n, rho, ndim, seed=11 + n)
    p = S.lj_params(n) if kind == O.LJ_ATTRACT_REPULSE else S.bidisperse_repulsion(n, n)
    masses = np.ones(n)
    masses[3] = np.inf  # a fixed atom
    v = S.gaussian_velocities(n, ndim, 1.0, seed=n)
    skin = 0.4 if kind == O.LJ_ATTRACT_REPULSE else 0.2
    dev = _make(kind, x, L, p, skin, masses=masses)
    dev.set_state(v=v)
    ref = _oracle_system(kind, x, L, p, skin, masses)
    ref.set_state(x=x, v=v, a=np.zeros_like(x))
    dev.initialize()
    ref.initialize()
    dt = 0.005 if integrator == "verlet" else 0.01
    # chunks of 1, 2, 3, 4 steps: single-step calls run drift + force kernels,
    # longer calls the fused force+kick+drift kernel between them
    for step, k in enumerate((1, 2, 3, 4)):
        if integrator == "verlet":
            dev.verlet(dt, k)
            ref.verlet(dt, k)
        else:
            dev.damped(dt, 0.5, k)
            ref.damped(dt, 0.5, k)
        xd, vd, ad, fd = dev.state()
        xr, vr, ar, fr = ref.state()
        assert dev.which() == ref.which(), step
        assert np.abs(xd - xr).max() <= 1e-12 * max(1.0, np.abs(xr).max()), step
        assert np.abs(vd - vr).max() <= 1e-11 * max(1.0, np.abs(vr).max()), step
        assert np.abs(fd - fr).max() <= 1e-10 * max(1.0, np.abs(fr).max()), step
    fd_, ld_ = dev.pairs()
    fr_, lr_ = ref.pairs()
    assert np.array_equal(fd_, fr_) and np.array_equal(ld_, lr_)
    assert abs(dev.kinetic_energy() - ref.kinetic_energy()) <= 1e-12 * ref.kinetic_energy()


def test_skin_rebuild_schedule_matches_oracle():
    _device_or_skip()
    n = 1500
    x, L = S.jittered_lattice(n, 0.8, 3, seed=21)
    p = S.lj_params(n)
    v = S.gaussian_velocities(n, 3, 2.0, seed=3)
    dev = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.15)
    dev.set_state(v=v)
    ref = _oracle_system(O.LJ_ATTRACT_REPULSE, x, L, p, 0.15, np.ones(n))
    ref.set_state(x=x, v=v, a=np.zeros_like(x))
    dev.initialize()
    ref.initialize()
    whiches = []
    for k in (1, 1, 5, 13, 40):  # fused multi-step calls rebuild mid-call
        dev.verlet(0.004, k)
        ref.verlet(0.004, k)
        whiches.append((dev.which(), ref.which()))
    assert all(a == b for a, b in whiches), whiches
    assert whiches[-1][0] > 3  # several rebuilds happened
    xd, vd, _, _ = dev.state()
    xr, vr, _, _ = ref.state()
    assert np.abs(xd - xr).max() <= 1e-10 * np.abs(xr).max()
    fd_, ld_ = dev.pairs()
    fr_, lr_ = ref.pairs()
    assert np.array_equal(fd_, fr_) and np.array_equal(ld_, lr_)


@pytest.mark.parametrize("integrator", ["verlet", "damped"])
def test_list_update_order_with_zero_skin(integrator):
    """Skin 0 rebuilds after every step that moved anything, and the forces
    of the next step use that list: the reference's order (forces at x_{q+1}
    with the list tested at x_q, then update_list at x_{q+1}).  Hundreds of
    pairs cross the cutoff per step here, so a force evaluated with a list
    rebuilt at x_{q+1} itself would differ far beyond the tolerance."""
    _device_or_skip()
    n = 1000
    x, L = S.jittered_lattice(n, 0.8, 3, seed=77)
    p = S.lj_params(n)
    v = S.gaussian_velocities(n, 3, 1.0, seed=78)
    dev = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.0)
    dev.set_state(v=v)
    ref = _oracle_system(O.LJ_ATTRACT_REPULSE, x, L, p, 0.0, np.ones(n))
    ref.set_state(x=x, v=v, a=np.zeros_like(x))
    dev.initialize()
    ref.initialize()
    for k in (1, 3, 20):
        if integrator == "verlet":
            dev.verlet(0.005, k)
            ref.verlet(0.005, k)
        else:
            dev.damped(0.005, 0.5, k)
            ref.damped(0.005, 0.5, k)
        xd, vd, _, fd = dev.state()
        xr, vr, _, fr = ref.state()
        assert dev.which() == ref.which(), k
        assert np.abs(xd - xr).max() <= 1e-11 * np.abs(xr).max(), k
        assert np.abs(fd - fr).max() <= 1e-9 * np.abs(fr).max(), k
    assert dev.which() == 1 + 24  # the initial build + one per step


def _tile_case_params(pset, n):
    if pset == "lj":
        return S.lj_params(n)
    if pset == "bi":
        return S.bidisperse_repulsion(n, n)
    if pset == "mono":  # one repulsion type
        return dict(eps=[1.0], sigma=[1.2], third=[2.0], types=np.zeros(n, np.int32),
                    diam=np.full(n, 1.2))
    # two LJ types
    rng = np.random.default_rng(n)
    types = (rng.random(n) < 0.3).astype(np.int32)
    sigma = np.array([1.0, 0.9])
    return dict(eps=[[1.0, 0.8], [0.8, 0.5]], sigma=list(sigma), third=[2.5, 2.5],
                types=types, diam=sigma[types] * 2.5)


@pytest.mark.parametrize("kind,ndim,n,rho,skin,pset", [
    (O.LJ_ATTRACT_REPULSE, 3, 30000, 0.8, 0.3, "lj"),    # row tiles, normal stencils
    (O.LJ_ATTRACT_REPULSE, 3, 4096, 0.8, 0.3, "lj"),     # full-row stencils
    (O.REPULSION, 3, 20000, 0.65, 0.2, "bi"),            # few atoms per cell
    (O.REPULSION, 2, 20000, 1.0, 0.2, "bi"),
    (O.REPULSION, 3, 20000, 0.65, 0.2, "mono"),          # one type: compact staging
    (O.LJ_ATTRACT_REPULSE, 3, 20000, 0.8, 0.3, "lj2"),   # two types: typed slots
])
@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED])
def test_tiled_layout_matches_untiled(kind, ndim, n, rho, skin, pset, prec, monkeypatch):
    """The LDS-tiled list and force kernels against the untiled ones on the
    same state: identical pair sets, forces within summation-order rounding
    (FP64) or the mixed bar, and identical trajectories' rebuild counts."""
    _device_or_skip()
    print(f"tiled vs untiled kind={kind} ndim={ndim} n={n} pset={pset} prec={prec}",
          flush=True)
    x, L = S.jittered_lattice(n, rho, ndim, seed=n + 5)
    p = _tile_case_params(pset, n)
    v = S.gaussian_velocities(n, ndim, 1.0, seed=9)
    sys_ = {}
    for tiled in (1, 0):
        monkeypatch.setenv("PARM_TILED", str(tiled))
        s = _make(kind, x, L, p, skin, prec=prec)
        s.set_state(v=v)
        s.initialize()
        sys_[tiled] = s
    a, b = sys_[1], sys_[0]
    fa, la = a.pairs()
    fb, lb = b.pairs()
    assert np.array_equal(fa, fb) and np.array_equal(la, lb)
    ref = O.forces(x, L, _table(kind, p), fb, lb)
    for s in (a, b):
        d = np.sqrt(np.sum((s.forces() - ref["f"]) ** 2, axis=1))
        if prec == capi.FP64:
            assert np.all(d <= 1e-12 * ref["fabs"] + 1e-15)
        else:
            # FP32 pair arithmetic: 2e-5 of sum_j |f_ij|, with a floor of
            # 1e-6 eps/sigma for atoms whose contacts barely overlap (the
            # harmonic force eps (1 - R/sigma)/sigma cancels in FP32 there)
            assert np.all(d <= 2e-5 * ref["fabs"] + 1e-6), np.max(d - 2e-5 * ref["fabs"])
    ua, ub = a.pair_sums(), b.pair_sums()
    assert abs(ua["U"] - ref["U"]) <= (1e-12 if prec == capi.FP64 else 1e-5) * abs(ref["U"])
    assert ua["contacts"] == ub["contacts"]
    for s in (a, b):
        s.verlet(0.002, 30)
    assert a.which() == b.which()
    xa_, xb_ = a.positions(), b.positions()
    assert np.abs(xa_ - xb_).max() <= (1e-10 if prec == capi.FP64 else 1e-5) * L[0]


def test_mixed_precision_force_error():
    _device_or_skip()
    n = 4096
    x, L = S.jittered_lattice(n, 0.8, 3, seed=99, jitter=0.1)
    p = S.lj_params(n)
    s64 = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3)
    s32 = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3, prec=capi.MIXED)
    for s in (s64, s32):
        s.update_list(True)
        s.set_forces(False)
    first, last = O.nlist(x, L, p["diam"], 0.3)
    ref = O.forces(x, L, _table(O.LJ_ATTRACT_REPULSE, p), first, last)
    err = S.normalised_force_error(s32.forces(), ref["f"], ref["fabs"])
    assert err <= 2e-5, err
    u32 = s32.pair_sums()["U"]
    assert abs(u32 - ref["U"]) <= 1e-5 * abs(ref["U"])


def test_errors_are_reported():
    _device_or_skip()
    s = capi.System(3, np.ones(4))
    with pytest.raises(ValueError):
        s.set_box([1.0, -1.0, 1.0])
    with pytest.raises(ValueError):
        s.nlist_set(0.3, [1.0, 1.0], members=[1, 1])
    with pytest.raises(RuntimeError):
        s.set_forces()  # no box / list / interaction yet
    with pytest.raises(ValueError):
        s.damped(0.0, 1.0, 1)


def test_workload_push_apart_min_separation():
    _device_or_skip()
    from parm_amd import workloads
    x, v, L = workloads.lj_liquid(3000, rho=0.8, T=1.0, min_sep=0.8, seed=5)
    first, last = O.nlist(x, L, np.full(3000, 0.8), 0.0, brute=True)
    assert len(first) == 0  # no pair closer than 0.8
    assert np.abs(v.sum(0)).max() < 1e-9
    assert abs(np.sum(v * v) / (3 * 3000 - 3) - 1.0) < 1e-12


def test_fp64_energy_conservation_large():
    """NVE at N = 65,536, 500 steps, dt = 0.002: a size-independent property."""
    _device_or_skip()
    from parm_amd import workloads
    n = 65536
    x, v, L = workloads.lj_liquid(n, seed=9)
    s = _make(O.LJ_ATTRACT_REPULSE, x, L, S.lj_params(n), 0.3)
    s.set_state(v=v, a=np.zeros_like(x))
    s.initialize()
    s.verlet(0.002, 200)  # let the initial push-apart configuration relax
    e0 = s.pair_sums()["U"] + s.kinetic_energy()
    es = []
    for _ in range(5):
        s.verlet(0.002, 100)
        es.append(s.pair_sums()["U"] + s.kinetic_energy())
    drift = max(abs(e - e0) for e in es) / abs(e0)
    assert drift < 1e-3, drift
    p, m = s.momentum()
    assert np.abs(p).max() < 1e-9 * n


@pytest.mark.parametrize("tpa", ["1", "2", "4", "8"])
def test_every_lanes_per_atom_layout(tpa, monkeypatch):
    """The force kernel's lanes-per-atom choices (autotuned at run time) all
    reproduce the oracle's forces and pair set."""
    _device_or_skip()
    monkeypatch.setenv("PARM_TPA", tpa)
    n = 3000
    x, L = S.jittered_lattice(n, 0.8, 3, seed=int(tpa))
    p = S.lj_params(n)
    s = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3)
    s.update_list(True)
    s.set_forces(False)
    first, last = O.nlist(x, L, p["diam"], 0.3)
    ref = O.forces(x, L, _table(O.LJ_ATTRACT_REPULSE, p), first, last)
    S.assert_fp64_forces(s.forces(), ref, x, L, _table(O.LJ_ATTRACT_REPULSE, p), first, last)
    f_dev, l_dev = s.pairs()
    assert np.array_equal(f_dev, first) and np.array_equal(l_dev, last)
    assert abs(s.pair_sums()["U"] - ref["U"]) <= 1e-12 * abs(ref["U"])


def test_rebuild_after_denser_packing_replans():
    """A rebuild whose largest tile stencil outgrows the previous plan by more
    than the speculative build's margin (the atoms squeezed into 3/4 of the
    box along y: a third more atoms per stencil row), so the build first
    queued against the old bound is redone from the plan as read; then the
    way back (a bound larger than needed).  Pair set and FP64 forces match
    the oracle each time."""
    _device_or_skip()
    n = 32768
    x, L = S.jittered_lattice(n, 0.8, 3, seed=11)
    p = S.lj_params(n)
    t = _table(O.LJ_ATTRACT_REPULSE, p)
    s = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3)
    squeezed = x.copy()
    squeezed[:, 1] = (squeezed[:, 1] + L[1] / 2) * 0.75 - L[1] / 2
    for step, xs in enumerate([x, squeezed, x]):
        print(f"replan case step {step}", flush=True)
        s.set_state(x=xs)
        assert s.update_list(True)
        first, last = O.nlist(xs, L, p["diam"], 0.3)
        f_dev, l_dev = s.pairs()
        assert np.array_equal(f_dev, first) and np.array_equal(l_dev, last)
        s.set_forces(False)
        ref = O.forces(xs, L, t, first, last)
        S.assert_fp64_forces(s.forces(), ref, xs, L, t, first, last)


def test_list_capacity_overflow_regrows():
    """Half the atoms squeezed into a quarter of the box (twice the density
    there): the fullest lists exceed the per-atom capacity the first build
    takes from the mean density (1.35 x the mean count + 24), so that build
    stores only the entries below the capacity, reports the counts, and is
    repeated with a larger capacity.  Pair set and FP64 forces match the
    oracle."""
    _device_or_skip()
    n = 32768
    x, L = S.jittered_lattice(n, 0.8, 3, seed=5)
    p = S.lj_params(n)
    t = _table(O.LJ_ATTRACT_REPULSE, p)
    xs = x - L * np.floor(x / L + 0.5)  # into [-L/2, L/2)
    left = xs[:, 0] < 0.0
    xs[left, 0] *= 0.5
    first, last = O.nlist(xs, L, p["diam"], 0.3)
    counts = np.bincount(np.concatenate([first, last]), minlength=n)
    r = p["diam"].max() + 0.3
    cap0 = 1.35 * n / np.prod(L) * (4.0 / 3.0 * np.pi * r**3) + 24.0
    assert counts.max() > cap0 + 10
    print(f"capacity overflow: max count {counts.max()}, first capacity {cap0:.0f}", flush=True)
    s = _make(O.LJ_ATTRACT_REPULSE, xs, L, p, 0.3)
    assert s.update_list(True)
    f_dev, l_dev = s.pairs()
    assert np.array_equal(f_dev, first) and np.array_equal(l_dev, last)
    s.set_forces(False)
    ref = O.forces(xs, L, t, first, last)
    S.assert_fp64_forces(s.forces(), ref, xs, L, t, first, last)


@pytest.mark.parametrize("case", ["empty", "no_members", "single"])
def test_empty_and_memberless_systems(case):
    """An AtomVec with no atoms, atoms none of which is a list member, and a
    single member atom: the step loops (long enough to run the autotuner),
    rebuild counting, energies and pair counts behave as the oracle's (no
    pairs, no forces, free flight; the lone atom's top-2 displacement is
    (d, 0), so it rebuilds once d reaches the skin)."""
    _device_or_skip()
    print(f"edge case {case}", flush=True)
    n = {"empty": 0, "no_members": 5, "single": 1}[case]
    L = np.full(3, 10.0)
    x = np.random.default_rng(0).uniform(-5, 5, size=(n, 3))
    v = np.full((n, 3), 0.1)
    members = np.zeros(0, np.int32) if case == "no_members" else None
    p = S.lj_params(n)
    s = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3, members=members)
    s.set_state(v=v, a=np.zeros_like(x))
    t = _table(O.LJ_ATTRACT_REPULSE, p)
    c = O.System(np.ones(n), L, t, p["diam"], 0.3, member=members)
    c.set_state(x=x, v=v, a=np.zeros_like(x))
    s.initialize()
    c.initialize()
    for kind in ("verlet", "damped"):
        if kind == "verlet":
            s.verlet(0.01, 20)
            c.verlet(0.01, 20)
        else:
            s.damped(0.01, 0.5, 20)
            c.damped(0.01, 0.5, 20)
        assert s.which() == c.which()
        assert s.numpairs() == 0
        assert s.pair_sums()["U"] == 0.0
        xs = s.positions()
        assert xs.shape == (n, 3)
        assert np.allclose(xs, c.state()[0], rtol=0, atol=1e-12)


@pytest.mark.parametrize("integrator", ["verlet", "damped"])
def test_nonfinite_position_is_reported(integrator):
    """A velocity set to +inf from the host drives one position to +inf in
    the first half step; the step loop reports PARM_E_NONFINITE (every
    kernel index comes from the list or is clamped, so the kernels run to
    completion on the non-finite values)."""
    _device_or_skip()
    n = 2000
    x, L = S.jittered_lattice(n, 0.8, 3, seed=21)
    p = S.lj_params(n)
    s = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3)
    v = np.zeros_like(x)
    v[7, 1] = np.inf
    s.set_state(v=v, a=np.zeros_like(x))
    s.initialize()
    with pytest.raises(capi.ParmError) as e:
        if integrator == "verlet":
            s.verlet(0.002, 30)
        else:
            s.damped(0.002, 0.5, 30)
    assert e.value.code == capi.E_NONFINITE, str(e.value)


@pytest.mark.parametrize("every", [1, 4])
def test_profile_samples_every_kth_step_launch(every, monkeypatch):
    """Instrumentation: with stride k the fused step launches of a run carry
    timing events at launches 0, k, 2k, ... (a 20-step run makes 19 fused
    launches; the last step is a plain force + kick), every rebuild carries
    them, and the trajectory is the same with and without them (the list
    layout pinned, so the autotuner cannot pick differently)."""
    _device_or_skip()
    monkeypatch.setenv("PARM_TPA", "4")
    monkeypatch.setenv("PARM_TILED", "1")
    n = 4096
    x, L = S.jittered_lattice(n, 0.8, 3, seed=31)
    p = S.lj_params(n)
    v = S.gaussian_velocities(n, 3, 1.0, seed=31)
    runs = []
    for prof in (True, False):
        s = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3)
        s.set_state(v=v, a=np.zeros_like(x))
        s.initialize()
        s.profile_enable(prof, every=every)
        w0 = s.which()
        s.verlet(0.002, 20)
        r = s.profile_read(reset=True)
        runs.append((s.state()[0], s.which() - w0, r))
    (x1, nreb, r), (x2, nreb2, r2) = runs
    assert r["step_launches"] == len(range(0, 19, every))
    assert r["step_ms"] > 0.0
    assert r["rebuilds"] == nreb == nreb2
    assert r2["step_launches"] == 0 and r2["rebuilds"] == 0
    assert np.array_equal(x1, x2)
    with pytest.raises(capi.ParmValueError):
        s.profile_enable(True, every=-1)


def test_skin_reduction_over_many_workgroups():
    """The skin test of 2.2 M atoms leaves more than 8192 workgroup partials,
    which several workgroups reduce with a last-workgroup merge: the two
    largest displacements, of two atoms far apart in the atom order, decide
    the rebuild at the reference's threshold (d1 + d2 >= skin,
    src/trackers.cpp:23-53), and a non-finite position is reported."""
    _device_or_skip()
    n = 2_200_000
    rng = np.random.default_rng(12)
    L = np.full(3, (n / 0.05) ** (1.0 / 3.0))
    x = rng.uniform(-L / 2, L / 2, (n, 3))
    p = S.lj_params(n)
    s = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3)
    assert s.update_list(True)
    w0 = s.which()
    for d1, d2, expect in ((0.2, 0.0999, False), (0.1999, 0.0, False),
                           (0.2, 0.1001, True)):
        print(f"skin case d1={d1} d2={d2}", flush=True)
        xs = x.copy()
        xs[0, 0] += d1
        xs[n - 1, 1] -= d2
        s.set_state(x=xs)
        assert s.update_list(False) == expect, (d1, d2)
        assert s.which() == w0 + (1 if expect else 0)
    xs = x.copy()
    xs[n // 2, 2] = np.inf
    s.set_state(x=xs)
    with pytest.raises(capi.ParmError) as e:
        s.update_list(False)
    assert e.value.code == capi.E_NONFINITE


def test_lane_count_bound_refused_before_allocation(monkeypatch):
    """Per-lane list counts are 16-bit: 70,000 atoms in a cube of edge 1.6
    (every pair within the list radius 2.8) at one lane per atom need
    69,999 entries in a lane; the build that finds it reports
    PARM_E_UNSUPPORTED instead of growing the list past the bound."""
    _device_or_skip()
    monkeypatch.setenv("PARM_TPA", "1")
    n = 70_000
    rng = np.random.default_rng(3)
    L = np.full(3, 50.0)
    x = rng.uniform(-0.8, 0.8, (n, 3))
    p = S.lj_params(n)
    s = _make(O.LJ_ATTRACT_REPULSE, x, L, p, 0.3)
    with pytest.raises((capi.ParmError, capi.ParmValueError)) as e:
        s.update_list(True)
    assert e.value.code == capi.E_UNSUPPORTED, str(e.value)
