/*
 * parm_oracle.h -- CPU oracle for the ParM per-timestep MD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (wackywendell/parm, double precision, VEC2D / VEC3D builds) used
 * as the checker by tests/, by __graft_entry__.smoke() and by the CPU-baseline
 * leg of bench.py.  The product (libparm_hip.so and the parm_amd package)
 * never links, loads or calls it.
 *
 * Every function restates a reference function; the citation is on each
 * declaration (paths relative to the reference repository root).
 *
 * Conventions shared with the reference:
 *   - positions are never wrapped; periodicity lives only in the minimum
 *     image  diff(r1, r2) = remainder(r1 - r2, L)  per component
 *     (src/box.hpp:68-78, 97-103);
 *   - |r|^2 is summed as (x*x + y*y) + z*z, the order Eigen's vectorised
 *     redux uses for a 3-vector of doubles (src/trackers.cpp:27,66 call
 *     Vec::norm()).  This order is an assumption about the compiled
 *     reference: the scalar Eigen unroller would give x*x + (y*y + z*z).
 *   - pair lists are in reference order: for member i ascending, for member
 *     j < i ascending, pair (first = i, last = j)  (src/trackers.cpp:56-69).
 *
 * Pair interaction kinds (per-type parameter tables; a "type" is one distinct
 * per-atom parameter tuple):
 *   OR_LJ_ATTRACT_REPULSE  NListed<IEpsSigCutAtom, LJAttractRepulsePair>
 *       eps_tab : ntypes x ntypes epsilon rows (row = type of atom1)
 *       sig_tab : sigma per type;  p3_tab : sigcut per type
 *   OR_REPULSION           NListed<EpsSigExpAtom, RepulsionPair>
 *       eps_tab : eps per type; sig_tab : sigma per type; p3_tab : exponent
 *   OR_LJ_CUT              NListed<EpsSigCutAtom, LennardJonesCutPair>
 *       eps_tab : eps per type; sig_tab : sigma per type; p3_tab : sigcut
 *   OR_LJ_CUT_II           NListed<IEpsISigCutAtom, LennardJonesCutPair>
 *       eps_tab, sig_tab : ntypes x ntypes rows; p3_tab : sigcut per type
 *   OR_LJ_REPULSE          NListed<EpsSigAtom, LJRepulsePair>
 *       eps_tab : eps per type; sig_tab : sigma per type; p3_tab unused
 *   OR_REPULSION_II        NListed<IEpsISigExpAtom, RepulsionPair>
 *       eps_tab, sig_tab : ntypes x ntypes rows; p3_tab : exponent per type
 *
 * What pins each part (tests/test_oracle.py, tests/test_oracle_kinds.py):
 * exact rationals, closed forms, 50-digit gradients, brute-force lists and
 * the reference's own test assertions; for CollectionNLCG, equilibrium at
 * the goal pressure, the exact reset/descend steps, finite termination
 * of conjugate gradients on an exact quadratic, and the second step's beta
 * against Polak and Ribiere's coefficient from the reported forces on a
 * non-quadratic packing (where Fletcher and Reeves' coefficient, the same
 * without the -f_new.f_old term, differs by over 20 %).
 */
#ifndef PARM_ORACLE_H
#define PARM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_LJ_ATTRACT_REPULSE = 1,
    OR_REPULSION = 2,
    OR_LJ_CUT = 3,      /* NListed<EpsSigCutAtom, LennardJonesCutPair>      */
    OR_LJ_CUT_II = 4,   /* NListed<IEpsISigCutAtom, LennardJonesCutPair>    */
    OR_LJ_REPULSE = 5,  /* NListed<EpsSigAtom, LJRepulsePair>               */
    OR_REPULSION_II = 6 /* NListed<IEpsISigExpAtom, RepulsionPair>          */
};

/* remainder(dx, L): OriginBox::diff per component (src/box.hpp:68-78). */
double or_vec_mod1(double dx, double L);

/* One pair, atom1 of type t1 at r1, atom2 of type t2 at r2 in an OriginBox.
 * Writes the force on atom1 (ndim doubles) and returns the pair energy.
 * LJ: src/interaction.hpp:1251-1299; Repulsion: src/interaction.hpp:1528-1566. */
double or_pair(int ndim, int kind, int ntypes, const double *eps_tab,
               const double *sig_tab, const double *p3_tab, int t1, int t2,
               const double *r1, const double *r2, const double *L,
               double *f_out);

/* NeighborList::update_list(true) pair predicate over members, O(M^2) brute
 * force exactly as src/trackers.cpp:56-69.  member[k] = atom index of the k-th
 * added member; diam[k] = its diameter.  Writes up to cap pairs (first, last)
 * as atom indices; returns the pair count (may exceed cap: then call again). */
int64_t or_nlist_brute(int ndim, int64_t nmembers, const int32_t *member,
                       const double *diam, const double *x, const double *L,
                       double skin, int64_t cap, int32_t *first,
                       int32_t *last, int64_t nexcl, const int32_t *ea,
                       const int32_t *eb);

/* Same pair set and order as or_nlist_brute, in O(M) with a periodic cell
 * grid and the identical predicate.  Used to check the device list at sizes
 * where the O(M^2) loop is too slow.  Both skip the nexcl excluded pairs
 * (ea[k], eb[k]) as NeighborList::ignore does (src/trackers.cpp:64). */
int64_t or_nlist_cells(int ndim, int64_t nmembers, const int32_t *member,
                       const double *diam, const double *x, const double *L,
                       double skin, int64_t cap, int32_t *first,
                       int32_t *last, int64_t nexcl, const int32_t *ea,
                       const int32_t *eb);

/* The O(N) skin test of NeighborList::update_list(false)
 * (src/trackers.cpp:23-53): running top-2 of |x_i - lastloc_i| over members
 * in order, returns 1 iff the sum of the two largest reaches skin. */
int or_skin_check(int ndim, int64_t nmembers, const int32_t *member,
                  const double *x, const double *lastlocs, double skin);

/* NListed<A,P>::set_forces / energy / pressure / stress over a pair list in
 * list order (src/interaction.hpp:2153-2291), Newton-3 scatter
 *   f[first] += f_pair ; f[last] -= f_pair.
 * f (natoms x ndim) is ACCUMULATED into (the caller zeroes it, as
 * AtomGroup::reset_forces does).  fabs_sum (natoms, may be NULL) accumulates
 * sum_j |f_ij| per atom, the normaliser for force comparisons.  U, W
 * (sum r.f) and S (ndim x ndim, sum r f^T; may be NULL) are written.
 * Also counts contacts (E != 0) and overlaps (E > 0)
 * (src/interaction.hpp:2125-2151). */
void or_forces(int ndim, int64_t natoms, const double *x, const double *L,
               int kind, int ntypes, const double *eps_tab,
               const double *sig_tab, const double *p3_tab, const int32_t *type,
               int64_t npairs, const int32_t *first, const int32_t *last,
               double *f, double *fabs_sum, double *U, double *W, double *S,
               uint64_t *contacts, uint64_t *overlaps);

/* ---------------- whole-system restatement (AoS, reference layout) -------- */

/* Opaque oracle system: struct Atom {Vec x,v,a,f; flt m;} per atom
 * (src/box.hpp:234-249), OriginBox, one NeighborList, one NListed<A,P>. */
typedef struct or_system or_system;

or_system *or_sys_create(int ndim, int64_t natoms, const double *masses,
                         const double *L, int kind, int ntypes,
                         const double *eps_tab, const double *sig_tab,
                         const double *p3_tab, const int32_t *type,
                         int64_t nmembers, const int32_t *member,
                         const double *diam, double skin, int brute_rebuild);
void or_sys_destroy(or_system *s);
/* Copy state in/out as natoms x ndim row-major arrays (any pointer may be NULL). */
void or_sys_set_state(or_system *s, const double *x, const double *v,
                      const double *a, const double *f);
void or_sys_get_state(const or_system *s, double *x, double *v, double *a,
                      double *f);
/* NeighborList::ignore(a, b) / ignore_size() (src/trackers.hpp:190-204). */
void or_sys_ignore(or_system *s, int32_t a, int32_t b);
int64_t or_sys_ignore_size(const or_system *s);
/* NeighborList::update_list(force)  (src/trackers.cpp:19-85); returns 1 if
 * the list was rebuilt. */
int or_sys_update_list(or_system *s, int force);
uint32_t or_sys_which(const or_system *s);
int64_t or_sys_numpairs(const or_system *s);
int64_t or_sys_get_pairs(const or_system *s, int64_t cap, int32_t *first,
                         int32_t *last);
/* Collection::set_forces(constraints_and_a) (src/collection.cpp:159-179)
 * with the single NListed interaction. */
void or_sys_set_forces(or_system *s, int and_a);
/* Observables (src/collection.cpp:73-142, src/box.cpp:239-260, 401-411). */
double or_sys_potential_energy(or_system *s);
double or_sys_virial(or_system *s);
void or_sys_stress(or_system *s, double *S);
double or_sys_kinetic_energy(const or_system *s, const double *originv);
void or_sys_momentum(const or_system *s, double *p);
double or_sys_mass(const or_system *s);
double or_sys_temp(const or_system *s, int minuscomv);
/* Collection::initialize() for a collection whose only tracker is the list
 * (src/collection.cpp:13-19). */
void or_sys_initialize(or_system *s);
/* CollectionVerlet::timestep (src/collection.cpp:442-469), nsteps times. */
void or_sys_verlet(or_system *s, double dt, int64_t nsteps);
/* CollectionDamped::set_constants + timestep (src/collection.cpp:342-381). */
void or_damped_constants(double dt, double damping, double *c0, double *c1,
                         double *c2);
void or_sys_damped(or_system *s, double dt, double damping, int64_t nsteps);

/* CollectionNLCG / CollectionNLCGFixedL on the system (src/collection.hpp:
 * 400-488, src/collection.cpp:494-870).  create runs Collection::initialize. */
typedef struct or_nlcg or_nlcg;
or_nlcg *or_nlcg_create(or_system *s, double dt, double P0, double kappa,
                        double kmax, int secmax, double seceps, int fixedL);
void or_nlcg_destroy(or_nlcg *c);
void or_nlcg_set_limits(or_nlcg *c, double alphamax, double afrac, double dxmax,
                        double stepmax, double maxdV);
void or_nlcg_set_forces(or_system *s, or_nlcg *c, int constraints_and_a, int setV);
void or_nlcg_reset(or_system *s, or_nlcg *c);
void or_nlcg_timestep(or_system *s, or_nlcg *c, int64_t nsteps);
void or_nlcg_descend(or_system *s, or_nlcg *c);
/* hamiltonian, pressure, kinetic_energy (src/collection.cpp:571-606) */
void or_nlcg_observables(or_system *s, or_nlcg *c, double *H, double *P, double *K);
/* alpha, beta, betaused, dxsum, alphavmax, sec, k, vl, fl, al, Knew, maxdV */
void or_nlcg_info(const or_nlcg *c, double *out);
/* OriginBox::box_shape() of the system's box */
void or_sys_box(const or_system *s, double *L);

#ifdef __cplusplus
}
#endif
#endif
