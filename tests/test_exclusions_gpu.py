"""NeighborList::ignore on the device (src/trackers.hpp:190-193,
src/trackers.cpp:64), against the CPU oracle: ignored pairs never enter the
list, in every list layout and lanes-per-atom setting; an ignore marks the
list for a rebuild at the next update; the exclusions persist through the
rebuilds of a run and the trajectory matches the oracle's."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from parm_amd import capi
from tests import systems as S

pytestmark = pytest.mark.gpu


def _device_or_skip():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible: the -m gpu tests need an MI355X")


def _exclusions(x, L, diam, skin, rng, k):
    """k pairs that are in the list (so the exclusion matters), a few that
    are not, one self pair and one repeated pair."""
    first, last = O.nlist(x, L, diam, skin)
    pick = rng.choice(len(first), size=min(k, len(first)), replace=False)
    ex = [(int(first[p]), int(last[p])) for p in pick]
    n = len(x)
    for _ in range(5):
        a, b = rng.integers(0, n, 2)
        ex.append((int(a), int(b)))
    ex.append((3, 3))
    ex.append(ex[0][::-1])
    return np.array(ex, np.int32)


@pytest.mark.parametrize("ndim,n,tiled,tpa", [(3, 3000, "1", "4"), (3, 3000, "0", "2"),
                                              (2, 4000, "1", "1"), (2, 4000, "0", "8")])
def test_ignored_pairs_leave_the_list(monkeypatch, ndim, n, tiled, tpa):
    _device_or_skip()
    print(f"exclusions ndim={ndim} n={n} tiled={tiled} tpa={tpa}", flush=True)
    monkeypatch.setenv("PARM_TILED", tiled)
    monkeypatch.setenv("PARM_TPA", tpa)
    rng = np.random.default_rng(n + ndim)
    x, L = S.jittered_lattice(n, 0.9, ndim, seed=n + 1)
    p = S.bidisperse_repulsion(n, n)
    ex = _exclusions(x, L, p["diam"], 0.3, rng, 400)
    s = capi.System(ndim, np.ones(n))
    s.set_box(L)
    s.nlist_set(0.3, p["diam"])
    s.set_repulsion(p["eps"], p["sigma"], p["third"], p["types"])
    s.set_state(x=x)
    s.update_list(True)
    w0 = s.which()
    s.ignore(ex[:, 0], ex[:, 1])
    uniq = {(max(a, b), min(a, b)) for a, b in ex.tolist()}
    assert s.ignore_size() == len(uniq)
    # ignorechanged: the next update rebuilds although nothing moved
    assert s.update_list(False)
    assert s.which() == w0 + 1
    f_dev, l_dev = s.pairs()
    f_ref, l_ref = O.nlist(x, L, p["diam"], 0.3, ignore=ex)
    assert np.array_equal(f_dev, f_ref) and np.array_equal(l_dev, l_ref)
    # and nothing moved since: no further rebuild
    assert not s.update_list(False)
    s.set_forces(False)
    table = O.PairTable(O.REPULSION, p["eps"], p["sigma"], p["third"], p["types"])
    ref = O.forces(x, L, table, f_ref, l_ref)
    S.assert_fp64_forces(s.forces(), ref, x, L, table, f_ref, l_ref)
    s.close()


def test_exclusions_persist_through_a_run():
    _device_or_skip()
    n, ndim = 2000, 3
    rng = np.random.default_rng(9)
    x, L = S.jittered_lattice(n, 0.8, ndim, seed=21)
    p = S.lj_params(n)
    table = O.PairTable(O.LJ_ATTRACT_REPULSE, p["eps"], p["sigma"], p["third"], p["types"])
    v = S.gaussian_velocities(n, ndim, 1.0, seed=4)
    ex = _exclusions(x, L, p["diam"], 0.3, rng, 300)
    s = capi.System(ndim, np.ones(n))
    s.set_box(L)
    s.nlist_set(0.3, p["diam"])
    s.set_lj(p["eps"], p["sigma"], p["third"], p["types"])
    s.set_state(x=x, v=v, a=np.zeros_like(x))
    ref = O.System(np.ones(n), L, table, p["diam"], 0.3)
    ref.set_state(x=x, v=v, a=np.zeros_like(x))
    for a, b in ex.tolist():
        ref.ignore(a, b)
    s.ignore(ex[:, 0], ex[:, 1])
    assert s.ignore_size() == ref.ignore_size()
    s.initialize()
    ref.initialize()
    s.verlet(0.004, 60)
    ref.verlet(0.004, 60)
    assert s.which() == ref.which() and s.which() > 2
    f_dev, l_dev = s.pairs()
    f_ref, l_ref = ref.pairs()
    assert np.array_equal(f_dev, f_ref) and np.array_equal(l_dev, l_ref)
    xs, _, _, _ = s.state()
    xr, _, _, _ = ref.state()
    assert np.abs(xs - xr).max() <= 1e-11 * np.abs(xr).max()
    # ignoring inside a run: the list rebuilds at the end of the next step
    w = s.which()
    s.ignore([int(f_dev[0])], [int(l_dev[0])])
    ref.ignore(int(f_ref[0]), int(l_ref[0]))
    s.verlet(0.004, 1)
    ref.verlet(0.004, 1)
    assert s.which() == ref.which() == w + 1
    f_dev, l_dev = s.pairs()
    f_ref, l_ref = ref.pairs()
    assert np.array_equal(f_dev, f_ref) and np.array_equal(l_dev, l_ref)
    # a new list (parm_nlist_set) starts without exclusions
    s.nlist_set(0.3, p["diam"])
    assert s.ignore_size() == 0
    s.close()


def test_ignore_rejects_bad_indices():
    _device_or_skip()
    s = capi.System(3, np.ones(5))
    s.set_box([5.0] * 3)
    s.nlist_set(0.3, np.ones(5))
    with pytest.raises(ValueError):
        s.ignore([0], [5])
    s.close()
