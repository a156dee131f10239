"""The drop-in API (parm_amd.d3 / d2 and the C++ headers) on the GPU.

The Hertzian-sphere test checks, on the device path, the invariants of the
reference's only hot-path test (pyparm/tests.py:297-317) with a harness of
its own: per-step relative energy change < 1e-2, mean temperature within
10 % of 1, std(E)/mean(E) < 1e-2.  The other tests compare the drop-in path
with the CPU oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from parm_amd import capi, d2, d3
from tests import systems as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hertzian_spheres(n=12, phi=0.3, seed=131):
    """Twelve bidisperse harmonic spheres through the drop-in module: half of
    diameter 1.0 and half 1.4, mass = diameter^3, Repulsion with eps 1.2 and
    exponent 2, a cubic OriginBox at packing fraction phi, NeighborList skin
    0.4, velocity Verlet dt = 0.01.  Seeded uniform positions, Gaussian
    velocities scaled by 1/m, random initial forces (overwritten by the first
    force evaluation)."""
    sig = np.repeat([1.0, 1.4], [n // 2, n - n // 2])
    masses = sig ** 3
    side = float((np.pi / 6 * np.sum(sig ** 3) / phi) ** (1 / 3))
    rng = np.random.default_rng(seed)
    box = d3.OriginBox(side)
    atoms = d3.AtomVec(masses.tolist())
    rep = d3.Repulsion(box, atoms, 0.4)
    for atom, s in zip(atoms, sig):
        rep.add(d3.EpsSigExpAtom(atom, 1.2, float(s), 2.0))
    atoms.set_positions(rng.uniform(0.0, side, (n, 3)))
    atoms.set_velocities(rng.normal(size=(n, 3)) / masses[:, None])
    atoms.set_forces(rng.normal(size=(n, 3)))
    rep.neighbor_list().update_list(True)
    collec = d3.CollectionVerlet(box, atoms, 0.01, [rep], [rep.neighbor_list()], [])
    return collec


def test_hertzian_spheres_conserve_energy_dropin():
    """The invariants the reference checks on this system
    (pyparm/tests.py:297-317): after 1000 thermostatted equilibration steps,
    10^4 NVE steps with every per-step relative energy change below 1e-2, the
    mean temperature (sampled every 10 steps) within 10 % of 1 and
    std(E) / mean(E) below 1e-2."""
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    collec = _hertzian_spheres()
    collec.scale_velocities_to_temp(1.0)
    for _ in range(1000):
        collec.timestep()
        collec.scale_velocities_to_temp(1.0)
    e_prev = collec.energy()
    samples = []
    for step in range(1, 10001):
        collec.timestep()
        e = collec.energy()
        assert abs(e - e_prev) <= 1e-2 * abs(e_prev) + 1e-8, (step, e, e_prev)
        e_prev = e
        if step % 10 == 0:
            samples.append((e, collec.temp()))
    energies, temps = np.asarray(samples).T
    assert abs(temps.mean() - 1.0) <= 0.1
    assert energies.std() / energies.mean() <= 1e-2


def _lj_dropin(n, seed, prec=None):
    rng = np.random.default_rng(seed)
    L = (n / 0.8) ** (1 / 3)
    m = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    x = (g + 0.5) * (L / m) + rng.normal(scale=0.05, size=(n, 3))
    v = rng.normal(size=(n, 3))
    v -= v.mean(0)
    box = d3.OriginBox(L)
    atoms = d3.AtomVec([1.0] * n)
    if prec is not None:
        atoms.set_device_precision(prec)
    lj = d3.LJAttractRepulse(box, atoms, 0.3)
    atoms.set_positions(x)
    atoms.set_velocities(v)
    for a in atoms:
        lj.add(d3.IEpsSigCutAtom(a, [1.0], 0, 1.0, 2.5))
    return box, atoms, lj, x, v, L


def test_dropin_lj_matches_oracle():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    n = 1500
    box, atoms, lj, x, v, L = _lj_dropin(n, 3)
    nl = lj.neighbor_list()
    collec = d3.CollectionVerlet(box, atoms, 0.002, [lj], [nl])
    t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], np.zeros(n, np.int32))
    ref = O.System(np.ones(n), [L] * 3, t, np.full(n, 2.5), 0.3)
    ref.set_state(x=x, v=v, a=np.zeros_like(x))
    ref.initialize()
    assert abs(collec.potential_energy() - ref.potential_energy()) <= 1e-12 * abs(
        ref.potential_energy())
    for _ in range(25):
        collec.timestep()
    ref.verlet(0.002, 25)
    xr, vr, ar, fr = ref.state()
    assert np.abs(atoms.get_positions() - xr).max() < 1e-11
    assert np.abs(atoms.get_forces() - fr).max() < 1e-9 * np.abs(fr).max()
    assert nl.which() == ref.which()
    first, last = ref.pairs()
    assert nl.numpairs() == len(first)
    got = [(p.first().n(), p.last().n()) for p in nl]
    assert got == list(zip(first.tolist(), last.tolist()))
    with pytest.raises(IndexError):
        nl[nl.numpairs()]
    assert abs(collec.kinetic_energy() - ref.kinetic_energy()) <= 1e-12 * ref.kinetic_energy()
    assert abs(lj.pressure(box) - ref.virial()) <= 1e-11 * abs(ref.virial())
    S = lj.stress(box)
    assert np.abs(S - ref.stress()).max() <= 1e-11 * np.abs(ref.stress()).max()
    # host edits between steps are uploaded before the next step
    a0 = atoms[0]
    a0.v = [0.0, 0.0, 0.0]
    collec.timestep()
    ref_state = ref.state()
    ref_v = ref_state[1].copy()
    ref_v[0] = 0
    ref.set_state(v=ref_v)
    ref.verlet(0.002, 1)
    assert np.abs(atoms.get_positions() - ref.state()[0]).max() < 1e-11


def test_dropin_timesteps_and_mixed_precision():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    box, atoms, lj, x, v, L = _lj_dropin(3000, 4, prec=d3.MIXED)
    collec = d3.CollectionVerlet(box, atoms, 0.002, [lj], [lj.neighbor_list()])
    e0 = collec.energy()
    collec.timesteps(200)
    assert abs(collec.energy() - e0) < 1e-3 * abs(e0)


def test_dropin_damped_2d_relaxation():
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    n = 2000
    rng = np.random.default_rng(8)
    sig = np.where(rng.random(n) < 0.5, 1.0, 1.4)
    L = float(np.sqrt(np.sum(np.pi / 4 * sig ** 2) / 0.8))
    box = d2.OriginBox(L)
    atoms = d2.AtomVec([1.0] * n)
    rep = d2.Repulsion(box, atoms, 0.2)
    atoms.set_positions(rng.uniform(0, L, (n, 2)))
    for a, s in zip(atoms, sig):
        rep.add(d2.EpsSigExpAtom(a, 1.0, float(s), 2.0))
    collec = d2.CollectionDamped(box, atoms, 0.05, 1.0, [rep], [rep.neighbor_list()])
    u0 = rep.energy(box)
    collec.timesteps(500)
    assert rep.energy(box) < 0.5 * u0
    assert rep.overlaps(box) <= rep.contacts(box)


def test_cpp_ljatoms_example_runs(tmp_path):
    """The reference driver's shape (src/bin/LJatoms.cpp) on the device:
    NListed<EpsSigCutAtom, LennardJonesCutPair>, random insertion with the
    energy check, NVE Verlet, an xyz frame per outer step."""
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    exe = os.path.join(ROOT, "parm_amd", "host", "examples", "LJatoms")
    xyz = tmp_path / "LJatoms.xyz"
    r = subprocess.run([exe, "600", "5", "400", str(xyz)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("step")]
    E = [float(ln.split()[3]) for ln in lines]
    E0 = float([ln for ln in r.stdout.splitlines() if ln.startswith("E0")][0].split()[1])
    assert len(E) == 5
    assert max(abs(e - E0) for e in E) < 1e-3 * abs(E0)
    frames = xyz.read_text().split("\n")
    assert len(frames) - 1 == 6 * 602 and frames[0] == "600"
    assert all(ln.startswith("C\t") for ln in frames[2:602])


def test_dropin_nlcg_packer_flow():
    """packer.cpp's flow (src/bin/packer.cpp:60-147) through the drop-in
    module: CollectionNLCG built first, tracker and interaction added after,
    a pressure goal set (reset), timesteps until mechanical equilibrium at
    the goal; the host OriginBox follows the device's resizing."""
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    n = 256
    rng = np.random.default_rng(3)
    sig = np.where(np.arange(n) < n // 2, 1.0, 1.4)
    L0 = float(np.sqrt(np.sum(np.pi / 4 * sig ** 2) / 0.5))
    box = d2.OriginBox(L0)
    atoms = d2.AtomVec([1.0] * n)
    hertz = d2.Repulsion(box, atoms, 0.1)
    for a, sg in zip(atoms, sig):
        a.x = rng.uniform(0, L0, 2)
        hertz.add(d2.EpsSigExpAtom(a, 1.0, float(sg), 2.0))
    nl = hertz.neighbor_list()
    nl.update_list(True)
    collec = d2.CollectionNLCG(box, atoms, 0.1, 1e-2)
    collec.add_tracker(nl)
    collec.add_interaction(hertz)
    collec.set_max_step(1e-2)
    collec.set_pressure_goal(1e-2)
    H0 = collec.hamiltonian()
    for _ in range(100):
        collec.timesteps(100)
        if abs(collec.pressure() / 1e-2 - 1) < 1e-9 and \
                np.abs(atoms.get_forces()).max() < 1e-11:
            break
    assert abs(collec.pressure() / 1e-2 - 1) < 1e-9
    assert np.abs(atoms.get_forces()).max() < 1e-11
    assert collec.hamiltonian() < H0
    assert box.V() != L0 * L0
    assert collec.get_pressure_goal() == 1e-2 and collec.k >= 0


def test_dropin_further_pair_kinds_and_ignore():
    """LJRepulse, LJIICut and RepulsionII through the drop-in module against
    the oracle's energies, and NeighborList.ignore removing pairs."""
    if capi.device_count() < 1:
        pytest.fail("no GPU visible")
    n = 800
    x, L = S.jittered_lattice(n, 0.8, 3, seed=12, jitter=0.1)
    types = (np.arange(n) % 2).astype(np.int32)
    cases = [
        ("LJRepulse", lambda a, t: d3.EpsSigAtom(a, [1.0, 0.5][t], [1.0, 1.2][t]),
         O.PairTable(O.LJ_REPULSE, [1.0, 0.5], [1.0, 1.2], None, types),
         np.array([1.0, 1.2])[types]),
        ("LJIICut", lambda a, t: d3.IEpsISigCutAtom(a, [[1.0, 0.7], [0.7, 0.9]][t],
                                                    [[1.0, 1.1], [1.1, 1.2]][t], t, 2.5),
         O.PairTable(O.LJ_CUT_II, [[1.0, 0.7], [0.7, 0.9]], [[1.0, 1.1], [1.1, 1.2]],
                     [2.5, 2.5], types),
         np.array([1.1 * 2.5, 1.2 * 2.5])[types]),
        ("RepulsionII", lambda a, t: d3.IEpsISigExpAtom(a, [[1.0, 0.8], [0.8, 1.2]][t],
                                                        [[1.0, 1.2], [1.2, 1.4]][t], t, 2.0),
         O.PairTable(O.REPULSION_II, [[1.0, 0.8], [0.8, 1.2]], [[1.0, 1.2], [1.2, 1.4]],
                     [2.0, 2.0], types),
         np.array([1.2, 1.4])[types]),
    ]
    for name, make_atom, table, diam in cases:
        print(f"drop-in pair kind {name}", flush=True)
        box = d3.OriginBox(float(L[0]))
        atoms = d3.AtomVec([1.0] * n)
        inter = getattr(d3, name)(box, atoms, 0.3)
        atoms.set_positions(x)
        for a, t in zip(atoms, types):
            inter.add(make_atom(a, int(t)))
        nl = inter.neighbor_list()
        nl.update_list(True)
        first, last = O.nlist(x, L, diam, 0.3)
        assert nl.numpairs() == len(first)
        ref = O.forces(x, L, table, first, last)
        assert abs(inter.energy(box) - ref["U"]) <= 1e-12 * abs(ref["U"])
        # ignore the first listed pair: the next update drops it
        p0 = nl[0]
        nl.ignore(p0.first(), p0.last())
        assert nl.ignore_size() == 1
        assert nl.update_list(False)
        f2, l2 = O.nlist(x, L, diam, 0.3, ignore=[(p0.first().n(), p0.last().n())])
        assert nl.numpairs() == len(f2) == len(first) - 1
        got = [(p.first().n(), p.last().n()) for p in nl]
        assert got == list(zip(f2.tolist(), l2.tolist()))
