#!/usr/bin/env python3
"""Benchmark of the per-timestep MD hot path on one MI355X.

Workload (the headline configuration): 3D Lennard-Jones, ParM's
LJAttractRepulsePair form V = eps (1 - sigma^6/r^6)^2 (eps = sigma = 1, cutoff
2.5 sigma, energy-shifted), N = 1,048,576 atoms of mass 1 at number density
0.8, positions uniform random (seeded) pushed apart to >= 0.8 sigma,
Gaussian velocities at T = 1 with zero momentum, neighbour skin 0.3,
velocity Verlet dt = 0.002, NVE.  A "step" is one CollectionVerlet::timestep()
including the neighbour-list skin test and any rebuild.

The step kernel is autotuned first (timed on its own, outside the timed
region); then `--warmup` untimed steps, then exactly `--steps` timed FP64
steps (the value) bracketed by device synchronisation.  The same system is
then run in mixed precision (FP32 pair arithmetic, FP64 state) over the same
steps.  Untimed beside them: 10^4-step NVE runs in both precisions (after a
1000-step equilibration from the pushed-apart start) for the energy drift and
the steady-state throughput, and the per-atom force error of the mixed form.

Roofline: the fused step kernel (forces + kick + next drift + skin test) is
timed with HIP events on its own stream (the end event is recorded right
after the kernel, before its skin-test reduction launches), every
PROFILE_EVERY-th launch of the timed steps (an event pair costs the stream a
few microseconds, ~3 % of a step when every launch carries one); its
algorithmic bytes per launch are
N * (192 + 4 * nbar) with nbar the mean full-list length (FP64 state read and
written once, one int32 list index per entry, neighbour gathers counted as
cache hits).  Peak: 8 TB/s HBM3E.

Also in the same line: the 1000-atom LJ case the reference runs on its CPU
("lj1000", beside the oracle's full run of it), the 262,144-disk 2D packing
relaxation ("relax2d") and the batched packings ("batched"): 4096 cubic
systems with atom counts uniform in [64, 4096], bidisperse 1.0 / 1.4 harmonic
spheres at packing fraction 0.64 from uniform random positions, damped Verlet
dt = 0.01 (damping 1.0), 2000 steps, mixed precision, all systems in one
batch with graph-replayed steps; beside it the same systems run one after
another through the single-system path (a sample of them).

CPU baseline: the oracle's restatement of ParM's single-core CPU path
(AoS atoms, half list, O(N^2) rebuild), timed on this host on a bounded
sample (see the "sample" field).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from parm_amd import capi, workloads  # noqa: E402

METRIC = "atom-timesteps/s, 1M-atom LJ FP64 vs mixed FP32 at stated drift; packings/s"
HBM_PEAK_GBS = 8000.0

# HBM bytes per atom moved by one launch of the FP64 fused step kernel on the
# 1,048,576-atom LJ workload, by list layout: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes (KiB per dispatch), FETCH_SIZE scaled by 2.0 and
# WRITE_SIZE by 1.0, the factors that make the same counters read 1 GiB for a
# 1 GiB device copy (profiles/k_force_lj_1M_pmc.txt).
TRAFFIC_BYTES_PER_ATOM = {  # (tiled list, lanes per atom) -> bytes per atom
    (True, 1): 401.6, (True, 2): 401.0, (True, 4): 404.5, (True, 8): 416.0,
    (False, 1): 559.3, (False, 2): 546.2, (False, 4): 563.3, (False, 8): 571.9,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_lj(x, v, L, prec, skin, masses=None):
    n, ndim = x.shape
    s = capi.System(ndim, np.ones(n) if masses is None else masses, prec)
    s.set_box(L)
    s.nlist_set(skin, np.full(n, 2.5))
    s.set_lj([[1.0]], [1.0], [2.5], np.zeros(n, np.int32))
    s.set_state(x=x, v=v, a=np.zeros_like(x))
    s.initialize()
    return s


def total_energy(s):
    return s.pair_sums()["U"] + s.kinetic_energy()


PROFILE_EVERY = 4


def run_timed(s, dt, steps, warmup):
    """Autotunes the step kernel (timed on its own), runs `warmup` untimed
    steps, then times exactly `steps` steps bracketed by device
    synchronisation.  Returns wall seconds, the tuning seconds and the profile
    dict of the timed steps."""
    s.synchronize()
    t0 = time.perf_counter()
    s.autotune(dt)
    t_tune = time.perf_counter() - t0
    if warmup:
        s.verlet(dt, warmup)
    s.synchronize()
    s.profile_read(reset=True)
    s.profile_enable(True, every=PROFILE_EVERY)
    t0 = time.perf_counter()
    s.verlet(dt, steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    prof = s.profile_read(reset=True)
    s.profile_enable(False)
    return wall, t_tune, prof


def drift_run(x, v, L, prec, skin, dt, equil, steps, every):
    """NVE energy conservation: `equil` untimed steps from the initial state,
    then `steps` steps with the total energy sampled every `every` steps
    (sampling outside the timing).  Returns the drift after equilibration
    (max |E(t) - E(t0)| / |E(t0)|), the drift from step 0 and the run's own
    throughput."""
    s = make_lj(x, v, L, prec, skin)
    s.autotune(dt)
    E = [total_energy(s)]
    done = 0
    while done < equil:
        k = min(every, equil - done)
        s.verlet(dt, k)
        done += k
        E.append(total_energy(s))
    i0 = len(E) - 1
    w0 = s.which()
    wall = 0.0
    done = 0
    while done < steps:
        k = min(every, steps - done)
        s.synchronize()
        t0 = time.perf_counter()
        s.verlet(dt, k)
        s.synchronize()
        wall += time.perf_counter() - t0
        done += k
        E.append(total_energy(s))
    rebuilds = s.which() - w0
    s.close()
    E = np.asarray(E)
    e0 = E[i0]
    return {
        "equilibration_steps": equil, "steps": steps, "sample_every": every,
        "energy_drift_rel_max": float(np.max(np.abs(E[i0:] - e0)) / abs(e0)),
        "energy_drift_rel_max_from_step0": float(np.max(np.abs(E - E[0])) / abs(E[0])),
        "E_start": float(e0), "E_end": float(E[-1]),
        "value": len(x) * steps / wall, "ms_per_step": 1e3 * wall / steps,
        "steps_per_rebuild": steps / max(rebuilds, 1),
    }


def cpu_baseline(x, v, L, skin, dt, gpu_steps_per_rebuild):
    """Time the oracle (single core) on a bounded sample of the same workload.
    The rebuild cost is amortised over the steady-state steps per rebuild of
    the GPU's 10^4-step run (the same trajectory; the skin rule decides)."""
    from oracle import oracle as O

    n = len(x)
    t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], np.zeros(n, np.int32))
    sysc = O.System(np.ones(n), L, t, np.full(n, 2.5), skin, ndim=3, brute=False)
    sysc.set_state(x=x, v=v, a=np.zeros_like(x))
    sysc.initialize()  # cell-list build (the brute O(N^2) build is timed below)
    w0 = sysc.which()
    nsteps = 3
    t0 = time.perf_counter()
    sysc.verlet(dt, nsteps)
    t_step = (time.perf_counter() - t0) / nsteps
    assert sysc.which() == w0, "sample steps must not rebuild"
    del sysc
    # the reference's O(N^2) rebuild, timed at N = 16384 and scaled by N^2
    nr = 16384
    xs = x[:nr] * (nr / n) ** (1 / 3)
    Ls = L * (nr / n) ** (1 / 3)
    tr = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], np.zeros(nr, np.int32))
    small = O.System(np.ones(nr), Ls, tr, np.full(nr, 2.5), skin, ndim=3, brute=True)
    small.set_state(x=xs)
    t0 = time.perf_counter()
    small.update_list(True)
    t_reb_small = time.perf_counter() - t0
    t_reb = t_reb_small * (n / nr) ** 2
    per_step = t_step + t_reb / max(gpu_steps_per_rebuild, 1.0)
    return {
        "value": n / per_step,
        "unit": "atom-timesteps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle restatement of ParM's CPU path, 1 core: {nsteps} non-rebuild "
                   f"Verlet steps at N={n} measured ({t_step:.3f} s/step); O(N^2) rebuild "
                   f"measured at N={nr} ({t_reb_small:.3f} s) and extrapolated by N^2 to "
                   f"{t_reb:.0f} s, amortised over {gpu_steps_per_rebuild:.2f} steps per "
                   f"rebuild (the GPU's 10^4-step FP64 run)"),
    }


def packings(nsys, nmin, nmax, phi, seed):
    """Batched workload: nsys cubic boxes, atom counts uniform in [nmin, nmax],
    bidisperse diameters 1.0 / 1.4 (50/50), packing fraction phi, uniform
    random positions (seeded)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, nsys)
    types, xs, Ls = [], [], []
    for n in sizes:
        t = np.zeros(n, np.int32)
        t[rng.permutation(n)[: n // 2]] = 1
        sig = np.array([1.0, 1.4])[t]
        L = float((np.sum(np.pi / 6 * sig ** 3) / phi) ** (1 / 3))
        types.append(t)
        xs.append(rng.uniform(0, L, (n, 3)))
        Ls.append(L)
    return sizes, np.concatenate(types), np.concatenate(xs), np.asarray(Ls)


def batched(args):
    """Packings: all systems in one batch (graph-replayed steps) vs the same
    systems run one after another through the single-system path."""
    prec = capi.MIXED if args.batch_precision == "mixed" else capi.FP64
    sizes, types, x, Ls = packings(args.batch_systems, 64, 4096, 0.64, args.seed + 1)
    ntot = int(sizes.sum())
    log(f"bench: batched packings, {len(sizes)} systems, {ntot} atoms")
    eps, sig, expo = [1.0, 1.0], [1.0, 1.4], [2.0, 2.0]
    b = capi.Batch(3, sizes, np.repeat(Ls[:, None], 3, 1), args.batch_skin, prec)
    b.set_repulsion(eps, sig, expo, types)
    b.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    b.initialize()
    U0, _ = b.sums()
    # autotune, then graph capture (17 steps = 8 graph replays, leaving the
    # position buffers as captured)
    t0 = time.perf_counter()
    b.autotune()
    t_tune = time.perf_counter() - t0
    b.damped(args.batch_dt, args.batch_damping, 17)
    b.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    b.initialize()
    w0, _ = b.lists()
    t0 = time.perf_counter()
    b.damped(args.batch_dt, args.batch_damping, args.batch_steps)
    U1, _ = b.sums()  # synchronises
    wall = time.perf_counter() - t0
    w1, npairs = b.lists()
    sums_end = b.pair_sums()
    # roofline of the dominant kernel: the same fused force kernel launched
    # eagerly with HIP events around each launch, over a further stretch of
    # the same run (the graph replays launch the same kernels)
    b.profile_read(reset=True)
    b.profile_enable(True)
    b.damped(args.batch_dt, args.batch_damping, args.batch_profile_steps, graph=False)
    prof = b.profile_read(reset=True)
    b.profile_enable(False)
    _, npairs_p = b.lists()
    b.close()
    nbar = 2.0 * float(np.sum(npairs_p)) / ntot
    B = 192.0 + 4.0 * nbar
    k_us = 1e3 * prof["force_ms"] / max(prof["force_launches"], 1)
    achieved = ntot * B / (k_us * 1e-6) / 1e9
    value = ntot * args.batch_steps / wall
    log(f"bench: batched {value:.3e} atom-steps/s, {len(sizes) / wall:.1f} packings/s, "
        f"{1e3 * wall / args.batch_steps:.3f} ms/step")
    # sequential: the first k systems, one parm_system each, same steps
    k = min(args.batch_seq_sample, len(sizes))
    off = np.concatenate([[0], np.cumsum(sizes)])
    seq_wall, seq_atoms = 0.0, 0
    for q in range(k):
        sl = slice(off[q], off[q + 1])
        n = int(sizes[q])
        s = capi.System(3, np.ones(n), prec)
        s.set_box([Ls[q]] * 3)
        s.nlist_set(args.batch_skin, np.array(sig)[types[sl]])
        s.set_repulsion(eps, sig, expo, types[sl])
        s.set_state(x=x[sl], v=np.zeros((n, 3)), a=np.zeros((n, 3)))
        s.initialize()
        s.synchronize()
        t0 = time.perf_counter()
        s.damped(args.batch_dt, args.batch_damping, args.batch_steps)
        s.synchronize()
        seq_wall += time.perf_counter() - t0
        seq_atoms += n
        s.close()
    seq_value = seq_atoms * args.batch_steps / seq_wall
    log(f"bench: sequential {seq_value:.3e} atom-steps/s over {k} systems")
    return {
        "workload": "3D bidisperse harmonic packings (RepulsionPair, exponent 2), damped "
                    "Verlet from uniform random positions",
        "systems": int(len(sizes)), "natoms_total": ntot, "natoms_range": [64, 4096],
        "packing_fraction": 0.64, "dt": args.batch_dt, "damping": args.batch_damping,
        "skin": args.batch_skin, "steps": args.batch_steps,
        "precision": args.batch_precision,
        "value": value, "unit": "atom-timesteps/s",
        "packings_per_s": len(sizes) / wall, "ms_per_step": 1e3 * wall / args.batch_steps,
        "rebuilds_per_system_mean": float(np.mean(w1 - w0)),
        "energy_per_atom_start": float(U0.sum() / ntot),
        "energy_per_atom_end": float(U1.sum() / ntot),
        "sequential": {"value": seq_value, "unit": "atom-timesteps/s",
                       "sample": f"the first {k} systems, one parm_system each, "
                                 f"{args.batch_steps} steps each, run one after another"},
        "speedup_vs_sequential": value / seq_value,
        "autotune_s": t_tune,
        "virial_per_atom_end": float(sums_end["W"].sum() / ntot),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_batch_force (fused force + kick + drift, all systems)",
                     "kernel_avg_us": k_us, "neighbours_per_atom": nbar,
                     "bytes_per_atom_step": B,
                     "rebuild_avg_us": 1e3 * prof["rebuild_ms"] /
                     max(prof["rebuild_launches"], 1),
                     "sample": f"{prof['force_launches']} eager launches after the timed "
                               f"run, HIP events on the batch's stream"},
    }


def relax2d_cpu_baseline(x, types, L, skin, dt, damping, gpu_steps_per_rebuild, steps):
    """Oracle (single core) on the relaxation workload: a few non-rebuild damped
    steps at full N on the given configuration, plus the O(N^2) rebuild timed
    at N = 16384 and scaled by N^2, amortised by the GPU run's rebuild rate."""
    from oracle import oracle as O

    n = len(x)
    sig = np.array([1.0, 1.4])
    t = O.PairTable(O.REPULSION, [1.0, 1.0], sig, [2.0, 2.0], types)
    sysc = O.System(np.ones(n), L, t, sig[types], skin, ndim=2, brute=False)
    sysc.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
    sysc.initialize()
    w0 = sysc.which()
    nsteps = 5
    t0 = time.perf_counter()
    sysc.damped(dt, damping, nsteps)
    t_step = (time.perf_counter() - t0) / nsteps
    assert sysc.which() == w0, "sample steps must not rebuild"
    del sysc
    nr = 16384
    scale = (nr / n) ** 0.5
    ts = O.PairTable(O.REPULSION, [1.0, 1.0], sig, [2.0, 2.0], types[:nr])
    small = O.System(np.ones(nr), L * scale, ts, sig[types[:nr]], skin, ndim=2, brute=True)
    small.set_state(x=x[:nr] * scale)
    t0 = time.perf_counter()
    small.update_list(True)
    t_reb_small = time.perf_counter() - t0
    t_reb = t_reb_small * (n / nr) ** 2
    per_step = t_step + t_reb / max(gpu_steps_per_rebuild, 1.0)
    return {
        "value": n / per_step,
        "unit": "atom-timesteps/s",
        "cores": 1,
        "kind": "port",
        "relaxations_per_s": 1.0 / (per_step * steps),
        "sample": (f"oracle restatement of ParM's CPU path, 1 core: {nsteps} non-rebuild "
                   f"damped steps at N={n} on the relaxed configuration ({t_step:.3f} s/step); "
                   f"O(N^2) rebuild measured at N={nr} ({t_reb_small:.3f} s) and extrapolated "
                   f"by N^2 to {t_reb:.0f} s, amortised over the GPU run's "
                   f"{gpu_steps_per_rebuild:.1f} steps per rebuild; relaxations/s = 1 / "
                   f"({steps} GPU steps x s/step)"),
    }


def relax2d(args):
    """2D bidisperse soft-disk packing relaxation: N disks, 50/50 diameters 1.0
    and 1.4, harmonic RepulsionPair (eps = 1, exponent 2), square box at
    packing fraction 0.84, seeded uniform positions at rest, CollectionDamped
    dynamics until the mean |F| per disk falls below the threshold (checked
    every chunk of steps) or the step cap.  Timed from the first damped step
    to the converged check; the chunk checks are inside the timed region."""
    prec = capi.MIXED if args.relax_precision == "mixed" else capi.FP64
    n = args.relax_natoms
    x, types, L = workloads.bidisperse_packing(n, 0.84, args.seed + 2, 2)
    sig = np.array([1.0, 1.4])

    def fresh():
        s = capi.System(2, np.ones(n), prec)
        s.set_box(L)
        s.nlist_set(args.relax_skin, sig[types])
        s.set_repulsion([1.0, 1.0], sig, [2.0, 2.0], types)
        s.set_state(x=x, v=np.zeros_like(x), a=np.zeros_like(x))
        s.initialize()
        s.synchronize()
        return s

    # untimed warm-up relaxation segment: autotune, first rebuild allocations
    s = fresh()
    s.damped(args.relax_dt, args.relax_damping, args.relax_chunk)
    s.synchronize()
    s.close()

    s = fresh()
    # the step kernel tuned before the timed run (reported apart)
    t0 = time.perf_counter()
    s.autotune(args.relax_dt, damped=True)
    t_tune = time.perf_counter() - t0
    w0 = s.which()
    trace = []
    steps = 0
    fmean = s.force_stats()[1]
    t0 = time.perf_counter()
    while steps < args.relax_max_steps and fmean >= args.relax_ftol:
        s.damped(args.relax_dt, args.relax_damping, args.relax_chunk)
        steps += args.relax_chunk
        fmax, fmean = s.force_stats()
        trace.append((steps, fmean))
    wall = time.perf_counter() - t0
    for st, fm in trace:
        log(f"bench: relax2d step {st}: mean |F| {fm:.3e}")
    rebuilds = s.which() - w0
    sums = s.pair_sums()
    xr = s.positions()
    s.close()
    converged = bool(fmean < args.relax_ftol)
    out = {
        "metric": "atom_timesteps_per_s",
        "value": n * steps / wall,
        "unit": "atom-timesteps/s",
        "n_gpus": 1,
        "steps": steps,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "dtype": "f64 state, f32 pair forces" if args.relax_precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "2D bidisperse harmonic disk packing relaxation (RepulsionPair "
                               "exponent 2, eps 1, diameters 1.0/1.4 50/50), damped dynamics "
                               "from uniform random positions to mean |F| < ftol",
                   "natoms": n, "packing_fraction": 0.84, "box": float(L[0]),
                   "dt": args.relax_dt, "damping": args.relax_damping,
                   "skin": args.relax_skin, "ftol": args.relax_ftol,
                   "check_every": args.relax_chunk, "seed": args.seed + 2},
        "converged": converged,
        "wall_s": wall,
        "relaxations_per_s": (1.0 / wall) if converged else None,
        "autotune_s": t_tune,
        "steps_per_rebuild": steps / max(rebuilds, 1),
        "mean_force_end": fmean,
        "energy_end": sums["U"],
        "overlaps_end": sums["overlaps"],
    }
    if not args.no_cpu:
        log("bench: relax2d cpu baseline ...")
        out["cpu_baseline"] = relax2d_cpu_baseline(
            xr, types, L, args.relax_skin, args.relax_dt, args.relax_damping,
            steps / max(rebuilds, 1), steps)
    return out


def lj1000(args):
    """The reference's CPU-runnable LJ case: N = 1000, density 0.8, T = 1,
    cutoff 2.5, skin 0.4, velocity Verlet dt = 0.001, 1000 steps, FP64,
    rebuilds included.  The oracle (the restatement of ParM's single-core
    path, with its O(N^2) rebuild) runs the same 1000 steps from the same state
    for the CPU figure, measured in full; the two final states are compared."""
    from oracle import oracle as O

    n, steps, dt, skin = 1000, 1000, 0.001, 0.4
    x, v, L = workloads.lj_liquid(n, rho=0.8, T=1.0, min_sep=0.8, seed=args.seed, log=log)
    # untimed warm-up system (autotuner, first allocations)
    w = make_lj(x, v, L, capi.FP64, skin)
    w.verlet(dt, 64)
    w.synchronize()
    w.close()
    s = make_lj(x, v, L, capi.FP64, skin)
    # the step kernel tuned before the timed run (same pairs, state untouched;
    # reported apart)
    t0 = time.perf_counter()
    s.autotune(dt)
    t_tune = time.perf_counter() - t0
    E0 = total_energy(s)
    w0 = s.which()
    s.synchronize()
    t0 = time.perf_counter()
    s.verlet(dt, steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    rebuilds = s.which() - w0
    xg = s.positions()
    dE = abs(total_energy(s) - E0) / abs(E0)
    s.close()
    out = {
        "metric": "atom_timesteps_per_s",
        "value": n * steps / wall,
        "unit": "atom-timesteps/s",
        "n_gpus": 1,
        "steps": steps,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded uniform positions pushed apart to 0.8 sigma, Gaussian "
                "velocities at T=1)",
        "config": {"workload": "3D LJ (LJAttractRepulsePair) NVE velocity Verlet, plain FP64, "
                               "the reference's CPU-runnable case",
                   "natoms": n, "density": 0.8, "cutoff": 2.5, "skin": skin, "dt": dt,
                   "T0": 1.0, "seed": args.seed},
        "rebuilds": rebuilds,
        "autotune_s": t_tune,
        "energy_drift_rel": dE,
    }
    if not args.no_cpu:
        log("bench: lj1000 cpu baseline (full run) ...")
        t = O.PairTable(O.LJ_ATTRACT_REPULSE, [[1.0]], [1.0], [2.5], np.zeros(n, np.int32))
        c = O.System(np.ones(n), L, t, np.full(n, 2.5), skin, ndim=3, brute=True)
        c.set_state(x=x, v=v, a=np.zeros_like(x))
        c.initialize()
        cw0 = c.which()
        t0 = time.perf_counter()
        c.verlet(dt, steps)
        cwall = time.perf_counter() - t0
        xc = c.state()[0]
        out["cpu_baseline"] = {
            "value": n * steps / cwall,
            "unit": "atom-timesteps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"oracle restatement of ParM's CPU path, 1 core: the full {steps}-step "
                      f"run with its O(N^2) rebuilds, measured ({cwall:.2f} s)",
        }
        out["parity"] = {
            "rebuilds_gpu": rebuilds,
            "rebuilds_cpu": c.which() - cw0,
            "max_abs_dx_final": float(np.max(np.abs(xg - xc))),
        }
    return out


def host_cpu():
    """Model name of the host CPU and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def pinned_to_one_core(fn, *a, **kw):
    """Runs fn pinned to one core (the reference is single-core), restoring
    the affinity afterwards.  Returns (result, core)."""
    before = os.sched_getaffinity(0)
    core = min(before)
    os.sched_setaffinity(0, {core})
    try:
        return fn(*a, **kw), core
    finally:
        os.sched_setaffinity(0, before)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--natoms", type=int, default=1 << 20)
    ap.add_argument("--skin", type=float, default=0.3)
    ap.add_argument("--dt", type=float, default=0.002)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-drift", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--drift-equil", type=int, default=1000)
    ap.add_argument("--drift-every", type=int, default=100)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 1000-atom LJ and 2D relaxation workloads")
    ap.add_argument("--batch-systems", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--batch-dt", type=float, default=0.01)
    ap.add_argument("--batch-damping", type=float, default=1.0)
    ap.add_argument("--batch-skin", type=float, default=0.2)
    ap.add_argument("--batch-precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--batch-seq-sample", type=int, default=32)
    ap.add_argument("--batch-profile-steps", type=int, default=40)
    ap.add_argument("--workload", choices=["lj", "relax2d", "lj1000"], default="lj",
                    help="lj (default): every workload in one JSON line; relax2d / lj1000: "
                         "only that workload, as its own JSON line")
    ap.add_argument("--relax-natoms", type=int, default=262144)
    ap.add_argument("--relax-precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--relax-dt", type=float, default=0.3)
    ap.add_argument("--relax-damping", type=float, default=0.02)
    ap.add_argument("--relax-skin", type=float, default=0.4)
    ap.add_argument("--relax-ftol", type=float, default=1e-10)
    ap.add_argument("--relax-chunk", type=int, default=250)
    ap.add_argument("--relax-max-steps", type=int, default=100000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: --gpus {args.gpus} is not supported "
              "(this build measures one MI355X; multi-GPU sharding is design-only)",
              file=sys.stderr)
        sys.exit(2)
    if capi.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "relax2d":
        print(json.dumps(relax2d(args)), flush=True)
        return
    if args.workload == "lj1000":
        print(json.dumps(lj1000(args)), flush=True)
        return

    n = args.natoms
    log(f"bench: building LJ liquid N={n}")
    t0 = time.perf_counter()
    x, v, L = workloads.lj_liquid(n, rho=0.8, T=1.0, min_sep=0.8, seed=args.seed, log=log)
    log(f"bench: workload ready in {time.perf_counter() - t0:.1f} s, L={L[0]:.4f}")

    # ---------------- FP64 (the value): exactly --steps timed steps
    s = make_lj(x, v, L, capi.FP64, args.skin)
    E0 = total_energy(s)
    wall, t_tune, prof = run_timed(s, args.dt, args.steps, args.warmup)
    np_pairs = s.numpairs()
    nbar = 2.0 * np_pairs / n
    layout = s.list_info()
    ms_per_step = 1e3 * wall / args.steps
    value = n * args.steps / wall
    k_avg_ms = prof["step_ms"] / max(prof["step_launches"], 1)
    B = 192.0 + 4.0 * nbar
    achieved = n * B / (k_avg_ms * 1e-3) / 1e9
    key = (layout["tiled"], layout["lanes"])
    traffic = (TRAFFIC_BYTES_PER_ATOM[key] * n
               if n == 1 << 20 and args.skin == 0.3 and key in TRAFFIC_BYTES_PER_ATOM
               else None)
    log(f"bench: FP64 {value:.3e} atom-steps/s, {ms_per_step:.4f} ms/step, kernel "
        f"{k_avg_ms * 1e3:.1f} us ({'tiled' if layout['tiled'] else 'untiled'} list, "
        f"{layout['lanes']} lanes/atom), {prof['rebuilds']} rebuilds "
        f"({prof['rebuild_ms'] / max(prof['rebuilds'], 1):.3f} ms each), nbar {nbar:.1f}, "
        f"autotune {t_tune:.3f} s")
    s.close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform positions pushed apart to 0.8 sigma, "
                "Gaussian velocities at T=1)",
        "config": {"workload": "3D LJ (LJAttractRepulsePair) NVE velocity Verlet, "
                               "plain FP64", "natoms": n, "density": 0.8,
                   "cutoff": 2.5, "skin": args.skin, "dt": args.dt, "T0": 1.0,
                   "E0": E0, "neighbours_per_atom": nbar,
                   "rebuilds_in_timed_steps": prof["rebuilds"],
                   "autotune_s": t_tune, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("k_force_tiled" if layout["tiled"] else "k_force") +
                               " (fused force + kick + drift + skin test)",
                     "list_layout": "tiled" if layout["tiled"] else "untiled",
                     "lanes_per_atom": layout["lanes"],
                     "kernel_avg_us": k_avg_ms * 1e3,
                     "bytes_per_atom_step": B},
    }

    # ---------------- mixed precision (same system, same seed, same steps)
    if not args.no_mixed:
        sm = make_lj(x, v, L, capi.MIXED, args.skin)
        wall_m, t_tune_m, prof_m = run_timed(sm, args.dt, args.steps, args.warmup)
        value_m = n * args.steps / wall_m
        lay_m = sm.list_info()
        sm.close()
        # force error vs FP64 at the initial positions
        a = make_lj(x, v, L, capi.FP64, args.skin)
        b = make_lj(x, v, L, capi.MIXED, args.skin)
        fa, fb = a.forces(), b.forces()
        a.close()
        b.close()
        d = np.sqrt(np.sum((fa - fb) ** 2, axis=1))
        fn = np.sqrt(np.sum(fa * fa, axis=1))
        out["mixed"] = {
            "value": value_m, "ms_per_step": 1e3 * wall_m / args.steps,
            "list_layout": "tiled" if lay_m["tiled"] else "untiled",
            "lanes_per_atom": lay_m["lanes"],
            "speedup_vs_fp64": value_m / value,
            "autotune_s": t_tune_m,
            "force_err_abs_max": float(d.max()),
            "force_err_abs_rms": float(np.sqrt(np.mean(d * d))),
            "force_err_rel_rms": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(fn * fn))),
            "kernel_avg_us": 1e3 * prof_m["step_ms"] / max(prof_m["step_launches"], 1),
        }
        log(f"bench: mixed {value_m:.3e} atom-steps/s ({value_m / value:.2f}x)")

    # ---------------- NVE drift over 10^4 steps, both precisions (untimed legs)
    steady_spr = None
    if not args.no_drift:
        drift = {}
        for name, prec in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
            if name == "mixed" and args.no_mixed:
                continue
            log(f"bench: {name} drift run ({args.drift_equil} + {args.drift_steps} steps)")
            drift[name] = drift_run(x, v, L, prec, args.skin, args.dt, args.drift_equil,
                                    args.drift_steps, args.drift_every)
            log(f"bench: {name} drift {drift[name]['energy_drift_rel_max']:.2e}, "
                f"{drift[name]['value']:.3e} atom-steps/s over the long run")
        if "fp64" in drift and "mixed" in drift:
            drift["mixed_speedup_vs_fp64"] = drift["mixed"]["value"] / drift["fp64"]["value"]
        out["nve_drift"] = drift
        steady_spr = drift["fp64"]["steps_per_rebuild"]

    if not args.no_extra:
        log("bench: 1000-atom LJ (the reference's CPU-runnable case)")
        out["lj1000"] = lj1000(args)
        log("bench: 2D packing relaxation")
        out["relax2d"] = relax2d(args)

    if not args.no_batch:
        out["batched"] = batched(args)

    if not args.no_cpu:
        log("bench: timing the CPU baseline sample (one core)")
        spr = steady_spr if steady_spr else 7.0
        cb, core = pinned_to_one_core(cpu_baseline, x, v, L, args.skin, args.dt, spr)
        model, ncores = host_cpu()
        cb["host"] = f"{model}; pinned to core {core} of {ncores} available"
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
