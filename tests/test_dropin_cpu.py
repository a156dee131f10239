"""Host-side behaviour of the drop-in API (parm_amd.d2 / d3): names,
signatures, Vec <-> numpy conversion, OriginBox geometry and the reference's
error conventions.  Without a GPU every device call must raise."""
import numpy as np
import pytest

from oracle import oracle as O
from parm_amd import capi, d2, d3


def test_reference_names_exist():
    for mod in (d2, d3):
        for name in ("OriginBox", "AtomVec", "AtomID", "Atom", "IDPair", "NeighborList",
                     "EpsSigExpAtom", "Repulsion", "IEpsSigCutAtom", "LJAttractRepulse",
                     "Interaction", "StateTracker", "Collection", "CollectionVerlet",
                     "CollectionDamped", "seed", "rand_vec"):
            assert hasattr(mod, name), name
    assert d3.NDIM == 3 and d2.NDIM == 2


def test_originbox_diff_is_remainder():
    rng = np.random.default_rng(0)
    L = np.array([3.0, 4.5, 7.25])
    box = d3.OriginBox(L)
    assert box.V() == np.prod(L)
    assert abs(box.L() - L.mean()) < 1e-15
    for _ in range(200):
        a, b = rng.uniform(-20, 20, 3), rng.uniform(-20, 20, 3)
        r = box.diff(a, b)
        assert all(r[k] == O.vec_mod(a[k] - b[k], L[k]) for k in range(3))
    b2 = d2.OriginBox(5.0)
    assert np.array_equal(b2.diff([3.0, -2.6], [0.0, 0.0]), [-2.0, 2.4])
    d3.seed(3)
    for _ in range(50):
        p = box.rand_loc()
        assert np.all(np.abs(p) <= L / 2)


def test_atomvec_state_and_iteration():
    av = d3.AtomVec([1.0, 2.0, 3.0])
    assert len(av) == 3 and av.size() == 3
    for i, a in enumerate(av):
        a.x = [i, 2 * i, 3 * i]
        a.v = np.array([1.0, 0.0, -1.0]) * i
    assert np.array_equal(av[2].x, [2, 4, 6]) and av[1].m == 2.0
    assert np.array_equal(av.get_positions()[:, 1], [0, 2, 4])
    av.set_velocities(np.ones((3, 3)))
    assert np.array_equal(av[0].v, [1, 1, 1])
    with pytest.raises(IndexError):
        av[3]
    av2 = d3.AtomVec(5, 1.5)
    assert av2[4].m == 1.5 and np.array_equal(av2[4].x, [0, 0, 0])


def test_duplicate_add_raises_value_error():
    box = d3.OriginBox(10.0)
    av = d3.AtomVec([1.0] * 3)
    nl = d3.NeighborList(box, av, 0.3)
    nl.add(av[0], 1.0)
    with pytest.raises(ValueError):
        nl.add(av[0], 1.0)
    assert nl.size() == 1


def test_one_list_per_atomvec_on_device_path():
    box = d3.OriginBox(10.0)
    av = d3.AtomVec([1.0] * 3)
    first = d3.NeighborList(box, av, 0.3)
    with pytest.raises(RuntimeError):
        d3.NeighborList(box, av, 0.3)
    del first
    d3.NeighborList(box, av, 0.3)  # released with its owner


def test_ignore_size_counts_unordered_pairs_once():
    """NeighborList::ignore_size (src/trackers.hpp:204) counts what
    PairList::add_pair stored (src/trackers.hpp:76-83): an unordered pair
    once, whichever order it was given in; a pair of an atom with itself is
    stored too.  Host-side bookkeeping, no device call."""
    box = d3.OriginBox(10.0)
    av = d3.AtomVec([1.0] * 4)
    nl = d3.NeighborList(box, av, 0.3)
    assert nl.ignore_size() == 0
    nl.ignore(av[0], av[1])
    nl.ignore(av[1], av[0])
    assert nl.ignore_size() == 1
    nl.ignore(av[2], av[2])
    nl.ignore(av[3], av[1])
    assert nl.ignore_size() == 3


def test_device_calls_raise_without_gpu():
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible here")
    box = d3.OriginBox(10.0)
    av = d3.AtomVec([1.0] * 4)
    r = d3.Repulsion(box, av, 0.4)
    for a in av:
        r.add(d3.EpsSigExpAtom(a, 1.0, 1.0, 2.0))
    with pytest.raises(RuntimeError, match="no HIP device"):
        r.energy(box)
    with pytest.raises(RuntimeError, match="no HIP device"):
        d3.CollectionVerlet(box, av, 0.01, [r], [r.neighbor_list()])
    with pytest.raises(ValueError):
        d3.CollectionDamped(box, av, -1.0, 1.0, [r], [r.neighbor_list()])


def test_atomvec_and_atom_pickle_round_trip():
    """Pickled state as the reference's SWIG extension stores it
    (src/sim.i:494-576): an Atom as (x, v, f, a); an AtomVec as
    (masses, [atom states]).  Host bookkeeping only, no device call."""
    import pickle

    av = d3.AtomVec([1.0, 2.0, 0.5])
    for i, a in enumerate(av):
        a.x = [i, 2.0 * i, -i]
        a.v = [0.5, i, 0.0]
        a.f = [1.0, 1.0, i]
        a.a = [0.0, -1.0, i]
    masses, states = av.__getstate__()
    assert masses == [1.0, 2.0, 0.5]
    assert states[1] == ((1.0, 2.0, -1.0), (0.5, 1.0, 0.0), (1.0, 1.0, 1.0), (0.0, -1.0, 1.0))
    assert av[2].__getstate__() == states[2]
    av2 = pickle.loads(pickle.dumps(av))
    assert av2.size() == 3 and [av2.getmass(i) for i in range(3)] == masses
    assert av2.__getstate__() == (masses, states)
    at = d3.Atom()
    at.x, at.v, at.f, at.a = [1, 2, 3], [4, 5, 6], [7, 8, 9], [0, 1, 0]
    at2 = pickle.loads(pickle.dumps(at))
    assert np.array_equal(at2.x, [1, 2, 3]) and np.array_equal(at2.a, [0, 1, 0])
    p2 = d2.AtomVec([1.0])
    p2[0].x = [0.25, -0.5]
    assert pickle.loads(pickle.dumps(p2)).get_positions().tolist() == [[0.25, -0.5]]
