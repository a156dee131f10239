"""CollectionNLCG / CollectionNLCGFixedL on the device (parm_nlcg_*),
against the oracle's restatement (src/collection.cpp:494-870): the first
timesteps follow the oracle's trajectory (positions, box, the scalar state),
differing only by the order of the device's reductions; run to convergence,
the device reaches the same mechanical equilibrium at the goal pressure."""
import numpy as np
import pytest

from oracle import oracle as O
from parm_amd import capi

pytestmark = pytest.mark.gpu


def _device_or_skip():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible: the -m gpu tests need an MI355X")


def _packing(n, ndim, phi, seed, prec=capi.FP64):
    rng = np.random.default_rng(seed)
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: n // 2]] = 1
    sig = np.array([1.0, 1.4])[types]
    Vs = np.sum(sig ** ndim) * (np.pi / 4 if ndim == 2 else np.pi / 6)
    Lside = (Vs / phi) ** (1 / ndim)
    x = rng.uniform(0, Lside, (n, ndim))
    t = O.PairTable(O.REPULSION, [1.0, 1.0], [1.0, 1.4], [2.0, 2.0], types)
    ref = O.System(np.ones(n), [Lside] * ndim, t, sig, 0.3, ndim=ndim)
    ref.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    s = capi.System(ndim, np.ones(n), prec)
    s.set_box([Lside] * ndim)
    s.nlist_set(0.3, sig)
    s.set_repulsion([1.0, 1.0], [1.0, 1.4], [2.0, 2.0], types)
    s.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    return s, ref


@pytest.mark.parametrize("ndim,n,fixed", [(2, 400, False), (3, 600, False), (2, 300, True)])
def test_nlcg_follows_the_oracle(ndim, n, fixed):
    _device_or_skip()
    print(f"nlcg ndim={ndim} n={n} fixed_L={fixed}", flush=True)
    s, ref = _packing(n, ndim, 0.6, seed=n + ndim)
    c = capi.NLCG(s, 0.1, 1e-2, fixed_L=fixed)
    cr = O.NLCG(ref, 0.1, 1e-2, fixed_L=fixed)
    for c_ in (c, cr):
        c_.set_limits(stepmax=1e-2)
    for step in range(5):
        c.timestep(1)
        cr.timestep(1)
        xd = s.positions()
        xr = ref.state()[0]
        scale = np.abs(xr).max()
        assert np.abs(xd - xr).max() <= 1e-9 * scale, (step, np.abs(xd - xr).max())
        assert np.allclose(s.box(), ref.box(), rtol=1e-12, atol=0)
        di, ri = c.info(), cr.info()
        assert di["sec"] == ri["sec"] and di["k"] == ri["k"], (step, di, ri)
        for key in ("alpha", "dxsum", "vl", "fl", "Knew"):
            assert abs(di[key] - ri[key]) <= 1e-8 * (abs(ri[key]) + 1e-12), (step, key)
    assert s.which() == ref.which()
    od, orf = c.observables(), cr.observables()
    for key in ("H", "P", "K"):
        assert abs(od[key] - orf[key]) <= 1e-8 * abs(orf[key]), key
    c.close()
    s.close()


@pytest.mark.parametrize("ndim", [2, 3])
def test_nlcg_reaches_equilibrium_on_device(ndim):
    _device_or_skip()
    s, _ = _packing(300 if ndim == 2 else 400, ndim, 0.55, seed=ndim)
    P0 = 1e-2
    c = capi.NLCG(s, 0.1, P0)
    c.set_limits(stepmax=1e-2)
    for _ in range(100):
        c.timestep(100)
        fmax = s.force_stats()[0]
        if fmax < 1e-11 and abs(c.observables()["P"] / P0 - 1) < 1e-9:
            break
    assert abs(c.observables()["P"] / P0 - 1) < 1e-9
    assert s.force_stats()[0] < 1e-11
    c.close()
    s.close()


def test_nlcg_reset_descend_and_params():
    _device_or_skip()
    s, ref = _packing(500, 3, 0.3, seed=5)
    c = capi.NLCG(s, 1e-3, 1.0)
    cr = O.NLCG(ref, 1e-3, 1.0)
    c.reset()
    cr.reset()
    _, v, a, f = s.state()
    assert np.array_equal(v, f) and np.array_equal(a, f)
    c.descend()
    cr.descend()
    assert np.allclose(s.box(), ref.box(), rtol=1e-14, atol=0)
    xr = ref.state()[0]
    assert np.abs(s.positions() - xr).max() <= 1e-12 * np.abs(xr).max()
    c.set_params(2e-3, 0.5, 5.0)
    c.reset()
    cr_info = c.info()
    assert cr_info["k"] == 0 and cr_info["vl"] == cr_info["al"] == cr_info["fl"]
    c.close()
    s.close()
