/*
 * parm_hip.h -- C ABI of libparm_hip.so: the MI355X (gfx950) device path for
 * ParM's per-timestep MD hot path.
 *
 * One `parm_system` is one ParM AtomVec (src/box.hpp:441-479) with its
 * OriginBox (src/box.hpp:97-157), one NeighborList (src/trackers.hpp:157-214)
 * and one neighbour-listed pair interaction (src/interaction.hpp:1876-1945),
 * one of the pair kinds of enum parm_pair_kind:
 *   NListed<IEpsSigCutAtom, LJAttractRepulsePair>  (SWIG name LJAttractRepulse)
 *   NListed<EpsSigExpAtom,  RepulsionPair>         (SWIG name Repulsion)
 *   NListed<EpsSigCutAtom,  LennardJonesCutPair>   (src/bin/LJatoms.cpp)
 *   NListed<IEpsISigCutAtom, LennardJonesCutPair>  (SWIG name LJIICut)
 *   NListed<EpsSigAtom,     LJRepulsePair>         (SWIG name LJRepulse)
 *   NListed<IEpsISigExpAtom, RepulsionPair>        (SWIG name RepulsionII)
 * (src/interaction.hpp:857-1018, 1251-1299, 1454-1590; src/sim.i:621-636),
 * driven by CollectionVerlet / CollectionDamped
 * (src/collection.hpp:273-295, 360-374).  All state lives on the current HIP
 * device at create time; every call is ordered on the system's own stream.
 *
 * Conventions
 *   - Arrays of vectors are natoms x ndim, row-major, double, in the caller's
 *     atom order (the AtomVec index).  The device keeps its own spatially
 *     sorted order internally; it is never visible through this ABI.
 *   - Every function returns 0 on success or a negative parm_status.  The
 *     message of the last failure on the calling thread is parm_last_error().
 *     Nothing throws across the ABI.
 *   - Precision PARM_FP64 computes everything in double; PARM_MIXED evaluates
 *     pair distances and forces in float, keeping positions, velocities, the
 *     integrator update and the energy/virial sums in double.
 *   - No function falls back to a CPU path: without a usable gfx950 device
 *     they fail with PARM_E_DEVICE.
 */
#ifndef PARM_HIP_H
#define PARM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct parm_system parm_system;

enum parm_status {
    PARM_OK = 0,
    PARM_E_INVALID = -1,   /* bad argument (maps to std::invalid_argument) */
    PARM_E_DEVICE = -2,    /* HIP runtime / device failure, or no GPU      */
    PARM_E_NOMEM = -3,     /* device allocation failed                    */
    PARM_E_UNSUPPORTED = -4, /* configuration the device path lacks        */
    PARM_E_NONFINITE = -5, /* a position became NaN or Inf                */
    PARM_E_STATE = -6      /* call out of order (e.g. no interaction set)  */
};

enum parm_precision { PARM_FP64 = 0, PARM_MIXED = 1 };

enum parm_pair_kind {
    PARM_PAIR_NONE = 0,
    PARM_PAIR_LJ_ATTRACT_REPULSE = 1, /* NListed<IEpsSigCutAtom, LJAttractRepulsePair>  */
    PARM_PAIR_REPULSION = 2,          /* NListed<EpsSigExpAtom, RepulsionPair>         */
    PARM_PAIR_LJ_CUT = 3,             /* NListed<EpsSigCutAtom, LennardJonesCutPair>   */
    PARM_PAIR_LJ_CUT_INDEXED = 4,     /* NListed<IEpsISigCutAtom, LennardJonesCutPair> */
    PARM_PAIR_LJ_REPULSE = 5,         /* NListed<EpsSigAtom, LJRepulsePair>            */
    PARM_PAIR_REPULSION_INDEXED = 6   /* NListed<IEpsISigExpAtom, RepulsionPair>       */
};

/* Message of the last failed call on this thread ("" if none). */
const char *parm_last_error(void);

/* ABI version: major changes break callers. */
int parm_version(int *major, int *minor);

/* Number of visible HIP devices (0 without a GPU; never an error). */
int parm_device_count(int *count);

/* ---------------------------------------------------------------- system */

/* AtomVec(vector<double> masses) / AtomVec(N, mass) (src/box.hpp:455-464):
 * natoms atoms with the given masses, x = v = a = f = 0. ndim is 2 or 3
 * (the reference's VEC2D / VEC3D builds, src/vecrand.hpp:3-16). */
int parm_system_create(int ndim, int precision, int64_t natoms,
                       const double *masses, parm_system **out);
int parm_system_destroy(parm_system *sys);
int parm_system_info(const parm_system *sys, int *ndim, int *precision,
                     int64_t *natoms);

/* OriginBox(Vec size) (src/box.hpp:102): edge lengths L[ndim] > 0. */
int parm_set_box(parm_system *sys, const double *L);
/* OriginBox::box_shape() (src/box.hpp:152): the current edges (the
 * minimiser below resizes the box). */
int parm_get_box(parm_system *sys, double *L);

/* Atom.x / .v / .a / .f assignment (src/box.hpp:234-249): any pointer may be
 * NULL to leave that field.  Uploads are ordered before later work. */
int parm_set_state(parm_system *sys, const double *x, const double *v,
                   const double *a, const double *f);
/* Read Atom.x / .v / .a / .f back (blocks until the device is done). */
int parm_get_state(parm_system *sys, double *x, double *v, double *a,
                   double *f);
/* Atom.m assignment. */
int parm_set_masses(parm_system *sys, const double *masses);

/* ----------------------------------------------------------- NeighborList */

/* NeighborList(box, atoms, skin) followed by add(atom_k, diameter_k) for
 * k = 0..nmembers-1 in order (src/trackers.cpp:10-17, trackers.hpp:194-201;
 * NListed::add passes A::max_size() as the diameter, interaction.hpp:1906).
 * member[k] is the atom index of the k-th added atom; duplicates are
 * rejected with PARM_E_INVALID like SubGroup::add (box.hpp:497-503).
 * Like add(), this marks the list for a rebuild at the next update. */
int parm_nlist_set(parm_system *sys, double skin, int64_t nmembers,
                   const int32_t *member, const double *diameter);

/* NeighborList::update_list(force) (src/trackers.cpp:19-85): the O(N) skin
 * test (two largest displacements since the last rebuild sum to >= skin),
 * then, if needed or forced, a rebuild.  *rebuilt (may be NULL) is 1 if the
 * list was rebuilt. */
int parm_nlist_update(parm_system *sys, int force, int *rebuilt);

/* NeighborList::ignore(a[k], b[k]) for k = 0..npairs-1 (src/trackers.hpp:
 * 190-193; PairList::add_pair :76-83): the unordered pairs of atom indices
 * never enter the list (the predicate skips them, src/trackers.cpp:64), and
 * the next update rebuilds (ignorechanged).  Pairs already ignored are kept
 * once.  parm_nlist_set starts a new list with no exclusions. */
int parm_nlist_ignore(parm_system *sys, int64_t npairs, const int32_t *a,
                      const int32_t *b);
/* NeighborList::ignore_size() (src/trackers.hpp:204): distinct ignored pairs. */
int parm_nlist_ignore_size(parm_system *sys, int64_t *n);

/* NeighborList::which() (rebuild counter) and numpairs() (trackers.hpp:187-188). */
int parm_nlist_which(parm_system *sys, uint32_t *which);
int parm_nlist_numpairs(parm_system *sys, int64_t *npairs);

/* NeighborList::get(i) / begin()..end() (trackers.hpp:205-212): the pair list
 * in the reference's order (member rank of `first` ascending, then of `last`
 * ascending, rank(last) < rank(first)) as atom indices.  Writes
 * min(cap, numpairs) pairs; *npairs receives numpairs. */
int parm_nlist_pairs(parm_system *sys, int64_t cap, int32_t *first,
                     int32_t *last, int64_t *npairs);

/* ------------------------------------------------------ pair interaction */

/* NListed<IEpsSigCutAtom, LJAttractRepulsePair> per-atom parameters, as a
 * type table: atom i has type atom_type[i] in [0, ntypes); type t carries
 * epsilons row eps_rows[t*ntypes .. +ntypes), indx t, sigma[t], sigcut[t]
 * (IEpsSigCutAtom(a, epsilons, indx, sigma, cut), interaction.hpp:997-1003).
 * Pair parameters follow LJAttractRepulsePair's constructor
 * (interaction.hpp:1256-1269).  ntypes <= 64.  The rows must be symmetric
 * (IEpsSigCutAtom::get_epsilon asserts it, interaction.hpp:1010-1016), else
 * PARM_E_UNSUPPORTED. */
int parm_pair_lj_attract_repulse(parm_system *sys, int ntypes,
                                 const double *eps_rows, const double *sigma,
                                 const double *sigcut,
                                 const int32_t *atom_type);

/* NListed<EpsSigExpAtom, RepulsionPair>: type t carries eps[t], sigma[t],
 * exponent[t] (EpsSigExpAtom(a, eps, sigma, exponent), interaction.hpp:1457).
 * ntypes <= 64. */
int parm_pair_repulsion(parm_system *sys, int ntypes, const double *eps,
                        const double *sigma, const double *exponent,
                        const int32_t *atom_type);

/* NListed<EpsSigCutAtom, LennardJonesCutPair> (the pair src/bin/LJatoms.cpp:
 * 37-38 uses): type t carries eps[t], sigma[t], sigcut[t]
 * (EpsSigCutAtom(a, epsilon, sigma, cut), interaction.hpp:897-905); pair
 * parameters eps = sqrt(e1 e2), sigma = (s1 + s2) / 2, cut = max(sigcut) and
 * the LennardJonesCut energy shift (interaction.hpp:238-281, 967-987).
 * List diameter of an atom: sigma * sigcut.  ntypes <= 64. */
int parm_pair_lj_cut(parm_system *sys, int ntypes, const double *eps,
                     const double *sigma, const double *sigcut,
                     const int32_t *atom_type);

/* NListed<IEpsISigCutAtom, LennardJonesCutPair> (SWIG name LJIICut,
 * src/sim.i:625): epsilons and sigmas as ntypes x ntypes rows, pair (t1, t2)
 * takes eps_rows[t1*ntypes+t2] and sigma_rows[t1*ntypes+t2]
 * (IEpsISigCutAtom::get_epsilon / get_sigma, interaction.hpp:911-956);
 * cut = max(sigcut).  List diameter: max of the atom's sigma row * sigcut.
 * Rows must be symmetric (each pair is evaluated from both atoms), else
 * PARM_E_UNSUPPORTED. */
int parm_pair_lj_cut_indexed(parm_system *sys, int ntypes, const double *eps_rows,
                             const double *sigma_rows, const double *sigcut,
                             const int32_t *atom_type);

/* NListed<EpsSigAtom, LJRepulsePair> (SWIG name LJRepulse, src/sim.i:621):
 * type t carries eps[t], sigma[t] (EpsSigAtom(a, epsilon, sigma),
 * interaction.hpp:857-866); eps = sqrt(e1 e2), sigma = (s1 + s2) / 2, the
 * purely repulsive LJ cut at sigma (LJRepulsive, interaction.hpp:119-147,
 * 875-891).  List diameter: sigma. */
int parm_pair_lj_repulse(parm_system *sys, int ntypes, const double *eps,
                         const double *sigma, const int32_t *atom_type);

/* NListed<IEpsISigExpAtom, RepulsionPair> (SWIG name RepulsionII,
 * src/sim.i:636): ntypes x ntypes eps and sigma rows as in
 * parm_pair_lj_cut_indexed, exponent[t] per type, n = (n1 + n2) / 2
 * (interaction.hpp:1477-1521, 1537-1542).  List diameter: max of the atom's
 * sigma row.  Rows must be symmetric. */
int parm_pair_repulsion_indexed(parm_system *sys, int ntypes, const double *eps_rows,
                                const double *sigma_rows, const double *exponent,
                                const int32_t *atom_type);

/* ------------------------------------------------------ forces, sums */

/* Collection::set_forces(and_a) with this one interaction
 * (src/collection.cpp:159-179): f = sum of pair forces over the current list;
 * if and_a, a = f/m (a = 0 for m <= 0 or m = inf). */
int parm_set_forces(parm_system *sys, int and_a);

/* NListed::set_forces (src/interaction.hpp:2165-2175): f += pair forces over
 * the current list (a is not touched). */
int parm_add_pair_forces(parm_system *sys);

/* One fused pass over the current list at the current positions:
 *   *U        NListed::energy            (interaction.hpp:2153-2163)
 *   *W        NListed::pressure (sum r.f) (interaction.hpp:2247-2261)
 *   S[ndim*ndim] NListed::stress (sum r f^T, row-major) (:2277-2291)
 *   *contacts, *overlaps  NListed::contacts / overlaps   (:2125-2151)
 * Any output may be NULL.  Forces are not modified. */
int parm_pair_sums(parm_system *sys, double *U, double *W, double *S,
                   uint64_t *contacts, uint64_t *overlaps);

/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The sums of parm_pair_sums, queued on the system's stream and
 * written to out, a DEVICE pointer to 13 doubles: U, W, S[9] (row-major
 * ndim x ndim in the first ndim*ndim), contacts, overlaps.  It neither
 * allocates nor synchronises (read out after parm_synchronize).  Its
 * workgroup partials live in a fixed device array of the library (about
 * 8 MiB per GPU, held until the process ends), so systems of at most
 * 2^21 atoms (larger: PARM_E_UNSUPPORTED), and calls that overlap on one GPU
 * race (order them on one stream, or synchronise between them). */
int parm_pair_sums_noalloc(parm_system *sys, double *out);

/* AtomGroup::kinetic_energy(originv) (box.cpp:401-411), momentum()
 * (box.cpp:251-260) and mass() (box.cpp:239-247).  originv may be NULL (0). */
int parm_kinetic_energy(parm_system *sys, const double *originv, double *K);
int parm_momentum(parm_system *sys, double *p, double *mass);

/* Largest and mean |f_i| over all atoms (relaxation criteria). */
int parm_force_stats(parm_system *sys, double *fmax, double *fmean);

/* ------------------------------------------------------------ integrators */

/* Collection::initialize() for a collection whose trackers are exactly this
 * system's NeighborList and whose interactions are this system's pair
 * (src/collection.cpp:13-19): update_list(false); set_forces(true);
 * update_list(false). */
int parm_initialize(parm_system *sys);

/* nsteps x CollectionVerlet::timestep() (src/collection.cpp:442-469),
 * including the NeighborList update at the end of each step. */
int parm_verlet_steps(parm_system *sys, double dt, int64_t nsteps);

/* nsteps x CollectionDamped::timestep() (src/collection.cpp:356-381) with the
 * constants of CollectionDamped::set_constants (:342-354). dt > 0. */
int parm_damped_steps(parm_system *sys, double dt, double damping,
                      int64_t nsteps);

/* Wait for all work queued on the system. */
int parm_synchronize(parm_system *sys);

/* No reference counterpart (performance tuning).  Picks the step kernel's
 * neighbour-list layout and lanes per atom now, by timing the candidates on
 * the current list (outputs redirected to scratch: the state is unchanged).
 * Without this call the first parm_verlet_steps / parm_damped_steps call of
 * >= 16 steps does it.  damped selects the kernel of that integrator. */
int parm_autotune(parm_system *sys, double dt, int damped);

/* ------------------------------------------------ nonlinear CG minimiser */
/* CollectionNLCG(box, atoms, dt, P0, {pair}, {list}, {}, kappa, kmax, secmax,
 * seceps) (src/collection.hpp:400-474, src/collection.cpp:494-523) on the
 * system: the pressure-controlled conjugate-gradient packing minimiser of
 * src/bin/packer.cpp:102 and pyparm/packmin.py:69, with the box resized in
 * place.  Creation runs Collection::initialize (src/collection.cpp:13-19).
 * fixed_L != 0 gives CollectionNLCGFixedL (only the first box edge changes,
 * src/collection.cpp:855-870).  The system must outlive the minimiser. */
typedef struct parm_nlcg parm_nlcg;
int parm_nlcg_create(parm_system *sys, double dt, double P0, double kappa, double kmax,
                     int secmax, double seceps, int fixed_L, parm_nlcg **out);
int parm_nlcg_destroy(parm_nlcg *c);
/* set_max_alpha, set_max_alpha_fraction, set_max_dx, set_max_step and the
 * maxdV member (src/collection.hpp:413-426, 473-477; defaults 2, 0, 100,
 * 1e-3, 0). */
int parm_nlcg_set_limits(parm_nlcg *c, double alphamax, double afrac, double dxmax,
                         double stepmax, double maxdV);
/* The parameter members dt, P0, kappa, kmax, secmax, seceps
 * (src/collection.hpp:408-421).  The reference's set_dt / set_pressure_goal /
 * set_kappa (:458-472) also call reset(): call parm_nlcg_reset after. */
int parm_nlcg_set_params(parm_nlcg *c, double dt, double P0, double kappa, double kmax,
                         int secmax, double seceps);
/* reset(), set_forces(constraints_and_a, setV), timestep() x nsteps and
 * descend() (src/collection.cpp:528-569, 685-853). */
int parm_nlcg_reset(parm_nlcg *c);
int parm_nlcg_set_forces(parm_nlcg *c, int constraints_and_a, int setV);
int parm_nlcg_timestep(parm_nlcg *c, int64_t nsteps);
int parm_nlcg_descend(parm_nlcg *c);
/* hamiltonian() = U + P0 V, pressure() = sum r.f / (NDIM V) and
 * kinetic_energy() (masses ignored) (src/collection.cpp:571-606); NULL skips. */
int parm_nlcg_observables(parm_nlcg *c, double *H, double *P, double *K);
/* The public tracking members, in this order: alpha, beta, betaused, dxsum,
 * alphavmax, sec, k, vl, fl, al, Knew, maxdV (12 doubles). */
int parm_nlcg_info(parm_nlcg *c, double *out);

/* ---------------------------------------------------- snapshots for output */
/* No reference counterpart as an interface: the state output of the
 * reference's drivers and Python pickling (src/bin/LJatoms.cpp:130-158,
 * src/sim.i:494-575) read Atom.x / .v on the host; here that copy leaves the
 * device without stalling the step loop.  begin queues, on the system's
 * stream, a gather of the positions (what bit 0) and / or velocities (bit 1)
 * into caller order, and a copy to pinned host memory on a stream of its own:
 * the call returns at once and later steps run while the copy travels.  The
 * last 4 snapshots stay readable by their ticket. */
int parm_snapshot_begin(parm_system *sys, int what, int64_t *ticket);
/* Waits for that snapshot and copies it out (natoms x ndim each; the box
 * edges at the time of the snapshot into L[ndim]); NULL skips. */
int parm_snapshot_read(parm_system *sys, int64_t ticket, double *x, double *v, double *L);
/* Writes that snapshot's positions as one xyz frame, as the reference's
 * drivers do: the atom count, the comment line, then per atom its element
 * letter (elements[i], or 'C' when elements is NULL) and the ndim components
 * of OriginBox::diff(0, x), tab-separated, printed like an ostream with the
 * given precision (6 is the default; src/bin/packer.cpp:170 uses 24).
 * append != 0 adds the frame to the file. */
int parm_snapshot_write_xyz(parm_system *sys, int64_t ticket, const char *path, int append,
                            const char *comment, const char *elements, int precision);

/* ------------------------------------------------------ instrumentation */
/* No reference counterpart (the reference has no timers); for benchmarks.
 * enable = k >= 1: every k-th fused step-kernel launch (force + kick + next
 * drift + skin test; k = 1 every one) and every list rebuild issued by
 * parm_verlet_steps / parm_damped_steps is bracketed by HIP events on the
 * system's stream (an event pair costs the stream a few microseconds);
 * enable = 0 turns it off.
 * parm_profile_read waits for the stream and returns the summed durations in
 * milliseconds and the counts since the last reset (reset != 0 clears them). */
int parm_profile_enable(parm_system *sys, int enable);
int parm_profile_read(parm_system *sys, int reset, double *step_ms,
                      int64_t *step_launches, double *rebuild_ms,
                      int64_t *rebuilds);
/* No reference counterpart (diagnostics).  The neighbour-list layout the step
 * kernels currently use: tiled = 1 for LDS-staged row tiles with 16-bit
 * entries, 0 for 32-bit atom-index entries; lanes = lanes per atom in the
 * force kernel; capacity = list entries per atom.  NULL skips. */
int parm_list_info(parm_system *sys, int *tiled, int *lanes, int *capacity);

/* ---------------------------------------------------------------- batches */
/* A parm_batch is nsys independent systems, each what the reference builds
 * with one AtomVec (src/box.hpp:441-479), one OriginBox (src/box.hpp:97-157),
 * one NeighborList with all atoms added (src/trackers.hpp:157-214), one
 * NListed pair interaction and one CollectionDamped / CollectionVerlet
 * (src/collection.hpp:273-295, 360-374).  The reference runs such systems one
 * after another; here every step covers all systems, each system keeps its
 * own box, list, skin test and rebuild count, and a run of steps needs no host
 * decision (it is replayed from a HIP graph).
 *
 * Atoms are concatenated system after system: system s owns atoms
 * [sum(natoms[0..s-1]), +natoms[s]).  Arrays are natoms_total x ndim.
 * Limits: at most 4096 list cells per system (larger boxes use larger cells),
 * list capacity fixed at parm_batch_pair_* time from the densest system; a run
 * that overflows it fails with PARM_E_STATE. */
typedef struct parm_batch parm_batch;

/* L: nsys x ndim box edges (OriginBox(Vec) per system, src/box.hpp:102);
 * skin: the NeighborList skin shared by all systems (src/trackers.cpp:10-17).
 * Masses start at 1; positions, velocities, accelerations at 0. */
int parm_batch_create(int ndim, int precision, int64_t nsys, const int64_t *natoms,
                      const double *L, double skin, parm_batch **out);
int parm_batch_destroy(parm_batch *b);
int parm_batch_info(const parm_batch *b, int *ndim, int *nsys, int64_t *natoms_total);
/* AtomVec masses (src/box.hpp:455-464); natoms_total values. */
int parm_batch_set_masses(parm_batch *b, const double *masses);
/* Atom.x / v / a / f of every atom (src/box.hpp:234-249); NULL skips. */
int parm_batch_set_state(parm_batch *b, const double *x, const double *v, const double *a);
int parm_batch_get_state(parm_batch *b, double *x, double *v, double *a, double *f);
/* NListed<EpsSigExpAtom, RepulsionPair>::add / NListed<IEpsSigCutAtom,
 * LJAttractRepulsePair>::add for every atom (src/interaction.hpp:1906-1910),
 * parameters per type as in parm_pair_repulsion / parm_pair_lj_attract_repulse;
 * atom_type has natoms_total entries.  Invalidates the lists. */
int parm_batch_pair_repulsion(parm_batch *b, int ntypes, const double *eps,
                              const double *sigma, const double *exponent,
                              const int32_t *atom_type);
int parm_batch_pair_lj_attract_repulse(parm_batch *b, int ntypes, const double *eps_rows,
                                       const double *sigma, const double *sigcut,
                                       const int32_t *atom_type);
/* Collection::initialize of every system (src/collection.cpp:13-19). */
int parm_batch_initialize(parm_batch *b);
/* nsteps of CollectionDamped::timestep (src/collection.cpp:342-381) /
 * CollectionVerlet::timestep (src/collection.cpp:442-469) on every system.
 * use_graph != 0 replays the steps from a captured HIP graph (same results). */
int parm_batch_damped_steps(parm_batch *b, double dt, double damping, int64_t nsteps,
                            int use_graph);
int parm_batch_verlet_steps(parm_batch *b, double dt, int64_t nsteps, int use_graph);
/* NListed::energy and NListed::contacts per system (src/interaction.hpp:
 * 2125-2150, 2153-2163); nsys values each, NULL skips. */
int parm_batch_sums(parm_batch *b, double *U, double *contacts);
/* Per system: NListed::energy (src/interaction.hpp:2153-2163), the virial
 * sum r.f of NListed::pressure (:2247-2261), NListed::contacts and overlaps
 * (:2125-2151).  nsys values each, NULL skips. */
int parm_batch_pair_sums(parm_batch *b, double *U, double *W, double *contacts,
                         double *overlaps);
/* No reference counterpart (performance tuning): picks the lanes per atom of
 * the batched force kernel now (otherwise the first run of >= 16 steps does). */
int parm_batch_autotune(parm_batch *b);
/* No reference counterpart (instrumentation).  While enabled, runs launch
 * their steps eagerly (same kernels as the graph replay) and bracket each
 * step's force launch and rebuild launch with HIP events; read returns the
 * summed milliseconds and counts (reset != 0 clears them). */
int parm_batch_profile_enable(parm_batch *b, int enable);
int parm_batch_profile_read(parm_batch *b, int reset, double *force_ms,
                            int64_t *force_launches, double *rebuild_ms,
                            int64_t *rebuild_launches);
/* NeighborList::which() and numpairs() per system (src/trackers.hpp:187-212). */
int parm_batch_lists(parm_batch *b, uint32_t *which, int64_t *npairs);
/* The pairs of system sysid's list as (first, last) atom indices within the
 * system, first > last, sorted by first then last.  Writes min(cap, n) pairs
 * and the total n. */
int parm_batch_pairs(parm_batch *b, int64_t sysid, int64_t cap, int32_t *first,
                     int32_t *last, int64_t *npairs);

#ifdef __cplusplus
}
#endif
#endif /* PARM_HIP_H */
