"""The further NListed pair kinds on the device, against the CPU oracle:

  NListed<EpsSigCutAtom, LennardJonesCutPair>    parm_pair_lj_cut
  NListed<IEpsISigCutAtom, LennardJonesCutPair>  parm_pair_lj_cut_indexed (LJIICut)
  NListed<EpsSigAtom, LJRepulsePair>             parm_pair_lj_repulse (LJRepulse)
  NListed<IEpsISigExpAtom, RepulsionPair>        parm_pair_repulsion_indexed (RepulsionII)

Bars as for the first two kinds: pair set bit-identical, FP64 forces within
1e-12 of sum_j |f_ij| per atom plus 1e-13 of the scale of the force's terms
(which cancel near the LJ minimum), U / W / S within 1e-12 relative, contacts
exact, 10-step Verlet trajectories within 1e-12 relative, mixed precision
forces within 2e-5 normalised plus 1e-6 of the terms' scale.
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


def _params(kind, n, seed):
    """Two atom types with distinct parameters, per the kind's atom struct;
    returns (oracle table args, device setter, list diameters)."""
    rng = np.random.default_rng(seed)
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: n // 2]] = 1
    if kind == O.LJ_CUT:
        eps, sig, cut = np.array([1.0, 0.7]), np.array([1.0, 1.1]), np.array([2.5, 2.2])
        diam = (sig * cut)[types]
        return (eps, sig, cut), lambda s: s.set_lj_cut(eps, sig, cut, types), diam, types
    if kind == O.LJ_CUT_II:
        eps = np.array([[1.0, 0.6], [0.6, 0.8]])
        sig = np.array([[1.0, 1.05], [1.05, 1.1]])
        cut = np.array([2.5, 2.3])
        diam = (sig.max(1) * cut)[types]
        return ((eps, sig, cut),
                lambda s: s.set_lj_cut_indexed(eps, sig, cut, types), diam, types)
    if kind == O.LJ_REPULSE:
        eps, sig = np.array([1.0, 1.5]), np.array([1.0, 1.2])
        diam = sig[types]
        return (eps, sig, None), lambda s: s.set_lj_repulse(eps, sig, types), diam, types
    eps = np.array([[1.0, 0.8], [0.8, 1.3]])
    sig = np.array([[1.0, 1.2], [1.2, 1.4]])
    expo = np.array([2.0, 2.5])
    diam = sig.max(1)[types]
    return ((eps, sig, expo),
            lambda s: s.set_repulsion_indexed(eps, sig, expo, types), diam, types)


CASES = [
    # kind, ndim, n, rho, skin
    (O.LJ_CUT, 3, 3000, 0.8, 0.3),
    (O.LJ_CUT, 2, 2500, 0.6, 0.3),
    (O.LJ_CUT_II, 3, 3000, 0.8, 0.3),
    (O.LJ_REPULSE, 3, 3000, 1.0, 0.3),
    (O.LJ_REPULSE, 2, 3000, 0.6, 0.2),
    (O.REPULSION_II, 3, 3000, 1.0, 0.2),
    (O.REPULSION_II, 2, 4000, 1.0, 0.2),
]


def _system(kind, ndim, n, rho, skin, seed, prec=capi.FP64):
    x, L = S.jittered_lattice(n, rho, ndim, seed=seed, jitter=0.12)
    (eps, sig, third), setter, diam, types = _params(kind, n, seed)
    s = capi.System(ndim, np.ones(n), prec)
    s.set_box(L)
    s.nlist_set(skin, diam)
    setter(s)
    s.set_state(x=x)
    table = O.PairTable(kind, eps, sig, third, types)
    return s, x, L, diam, table


@pytest.mark.parametrize("kind,ndim,n,rho,skin", CASES)
def test_pair_kind_forces_energy_and_pairs(kind, ndim, n, rho, skin):
    _device_or_skip()
    print(f"pair kind {kind} ndim={ndim} n={n}", flush=True)
    s, x, L, diam, table = _system(kind, ndim, n, rho, skin, seed=11 * n + kind)
    s.update_list(True)
    f_dev, l_dev = s.pairs()
    first, last = O.nlist(x, L, diam, skin)
    assert np.array_equal(f_dev, first) and np.array_equal(l_dev, last)
    s.set_forces(False)
    ref = O.forces(x, L, table, first, last)
    S.assert_fp64_forces(s.forces(), ref, x, L, table, first, last)
    sums = s.pair_sums()
    assert abs(sums["U"] - ref["U"]) <= 1e-12 * abs(ref["U"]) + 1e-13
    assert abs(sums["W"] - ref["W"]) <= 1e-12 * abs(ref["W"]) + 1e-13
    assert np.abs(sums["S"] - ref["S"]).max() <= 1e-12 * (np.abs(ref["S"]).max() + 1e-300)
    assert sums["contacts"] == ref["contacts"]
    assert sums["overlaps"] == ref["overlaps"]
    s.close()


@pytest.mark.parametrize("kind,ndim,n,rho,skin", CASES[::2])
def test_pair_kind_verlet_trajectory(kind, ndim, n, rho, skin):
    _device_or_skip()
    print(f"trajectory kind {kind} ndim={ndim} n={n}", flush=True)
    s, x, L, diam, table = _system(kind, ndim, n, rho, skin, seed=3 * n + kind)
    v = S.gaussian_velocities(n, ndim, 0.5, seed=n)
    s.set_state(v=v, a=np.zeros_like(x))
    s.initialize()
    ref = O.System(np.ones(n), L, table, diam, skin, ndim=ndim)
    ref.set_state(x=x, v=v, a=np.zeros_like(x))
    ref.initialize()
    s.verlet(0.002, 10)
    ref.verlet(0.002, 10)
    xs, vs, _, _ = s.state()
    xr, vr, _, _ = ref.state()
    assert np.abs(xs - xr).max() <= 1e-12 * np.abs(xr).max()
    assert np.abs(vs - vr).max() <= 1e-11 * (np.abs(vr).max() + 1.0)
    assert s.which() == ref.which()
    s.close()


@pytest.mark.parametrize("kind", [O.LJ_CUT, O.LJ_REPULSE, O.REPULSION_II])
def test_pair_kind_mixed_precision_forces(kind):
    """Mixed precision: FP32 pair arithmetic, so the per-atom error is bounded
    by 2e-5 of sum_j |f_ij| plus 1e-6 of the FP32-rounded terms' scale."""
    _device_or_skip()
    n, ndim = 4000, 3
    s, x, L, diam, table = _system(kind, ndim, n, 0.9, 0.3, seed=kind, prec=capi.MIXED)
    s.update_list(True)
    s.set_forces(False)
    first, last = O.nlist(x, L, diam, 0.3)
    ref = O.forces(x, L, table, first, last)
    d = np.sqrt(((s.forces() - ref["f"]) ** 2).sum(1))
    bar = 2e-5 * ref["fabs"] + 1e-6 * S.pair_term_scale(x, L, table, first, last)
    assert np.all(d <= bar), float(np.max(d / np.maximum(bar, 1e-300)))
    s.close()


def test_asymmetric_rows_are_refused():
    _device_or_skip()
    n = 10
    s = capi.System(3, np.ones(n))
    s.set_box([10.0] * 3)
    s.nlist_set(0.3, np.full(n, 1.4))
    t = np.array([0, 1] * 5, np.int32)
    with pytest.raises(capi.ParmError, match="symmetric"):
        s.set_repulsion_indexed([[1.0, 0.5], [0.6, 1.0]], [[1.0, 1.2], [1.2, 1.4]],
                                [2.0, 2.0], t)
    with pytest.raises(capi.ParmError, match="symmetric"):
        s.set_lj_cut_indexed([[1.0, 0.5], [0.5, 1.0]], [[1.0, 1.2], [1.1, 1.4]],
                             [2.5, 2.5], t)
    s.close()
