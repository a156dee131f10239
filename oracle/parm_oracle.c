/*
 * parm_oracle.c -- CPU oracle for the ParM per-timestep MD hot path.
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, smoke() and the
 * CPU-baseline leg of bench.py.  Never part of the product.  See
 * parm_oracle.h for the conventions and the reference citations.
 *
 * Built with -O2 -ffp-contract=off (no fused multiply-add), so every
 * expression rounds exactly as written, like the reference's -O2 x86-64
 * build (makefile:14) without -mfma.
 */
#include "parm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_system {
    int ndim;
    int64_t natoms;
    void *atoms;       /* AoS, reference layout */
    double L[3];
    int kind, ntypes;
    double *eps_tab, *sig_tab, *p3_tab;
    int32_t *type;     /* per atom */
    int64_t nmembers;
    int32_t *member;   /* atom index of each member, in add order */
    double *diam;      /* per member */
    double skin;
    int brute;
    double *lastlocs;  /* per member */
    double *xtmp;      /* natoms x ndim scratch */
    int32_t *first, *last;
    int64_t npairs, pcap;
    uint32_t updatenum;
    int ignorechanged;
    void *pairs;
    uint32_t pairs_which;
};

#define SUF(name) name##_2d
#define NDIM 2
#include "parm_oracle_impl.h"
#undef NDIM
#undef SUF
#define SUF(name) name##_3d
#define NDIM 3
#include "parm_oracle_impl.h"
#undef NDIM
#undef SUF

double or_vec_mod1(double dx, double L) { return remainder(dx, L); }

double or_pair(int ndim, int kind, int ntypes, const double *eps_tab,
               const double *sig_tab, const double *p3_tab, int t1, int t2,
               const double *r1, const double *r2, const double *L,
               double *f_out) {
    if (ndim == 3) {
        pairP_3d p;
        double r[3];
        pair_params_3d(kind, ntypes, eps_tab, sig_tab, p3_tab, t1, t2, &p);
        diff_3d(r1, r2, L, r);
        return pair_eval_3d(kind, &p, r, f_out);
    } else {
        pairP_2d p;
        double r[2];
        pair_params_2d(kind, ntypes, eps_tab, sig_tab, p3_tab, t1, t2, &p);
        diff_2d(r1, r2, L, r);
        return pair_eval_2d(kind, &p, r, f_out);
    }
}

int64_t or_nlist_brute(int ndim, int64_t M, const int32_t *member,
                       const double *diam, const double *x, const double *L,
                       double skin, int64_t cap, int32_t *first,
                       int32_t *last) {
    return ndim == 3
               ? nlist_brute_3d(M, member, diam, x, L, skin, cap, first, last)
               : nlist_brute_2d(M, member, diam, x, L, skin, cap, first, last);
}

int64_t or_nlist_cells(int ndim, int64_t M, const int32_t *member,
                       const double *diam, const double *x, const double *L,
                       double skin, int64_t cap, int32_t *first,
                       int32_t *last) {
    return ndim == 3
               ? nlist_cells_3d(M, member, diam, x, L, skin, cap, first, last)
               : nlist_cells_2d(M, member, diam, x, L, skin, cap, first, last);
}

int or_skin_check(int ndim, int64_t M, const int32_t *member, const double *x,
                  const double *lastlocs, double skin) {
    return ndim == 3 ? skin_check_3d(M, member, x, lastlocs, skin)
                     : skin_check_2d(M, member, x, lastlocs, skin);
}

void or_forces(int ndim, int64_t natoms, const double *x, const double *L,
               int kind, int ntypes, const double *eps_tab,
               const double *sig_tab, const double *p3_tab, const int32_t *type,
               int64_t npairs, const int32_t *first, const int32_t *last,
               double *f, double *fabs_sum, double *U, double *W, double *S,
               uint64_t *contacts, uint64_t *overlaps) {
    if (ndim == 3)
        forces_3d(natoms, x, L, kind, ntypes, eps_tab, sig_tab, p3_tab, type,
                  npairs, first, last, f, fabs_sum, U, W, S, contacts,
                  overlaps);
    else
        forces_2d(natoms, x, L, kind, ntypes, eps_tab, sig_tab, p3_tab, type,
                  npairs, first, last, f, fabs_sum, U, W, S, contacts,
                  overlaps);
}

static double *dup_d(const double *p, int64_t n) {
    double *q = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (n > 0) memcpy(q, p, sizeof(double) * n);
    return q;
}

or_system *or_sys_create(int ndim, int64_t natoms, const double *masses,
                         const double *L, int kind, int ntypes,
                         const double *eps_tab, const double *sig_tab,
                         const double *p3_tab, const int32_t *type,
                         int64_t nmembers, const int32_t *member,
                         const double *diam, double skin, int brute_rebuild) {
    or_system *s = (or_system *)calloc(1, sizeof(or_system));
    s->ndim = ndim;
    s->natoms = natoms;
    size_t asz = ndim == 3 ? sizeof(atom_3d) : sizeof(atom_2d);
    s->atoms = calloc((size_t)(natoms > 0 ? natoms : 1), asz);
    for (int64_t i = 0; i < natoms; i++) {
        if (ndim == 3)
            ((atom_3d *)s->atoms)[i].m = masses[i];
        else
            ((atom_2d *)s->atoms)[i].m = masses[i];
    }
    for (int k = 0; k < ndim; k++) s->L[k] = L[k];
    s->kind = kind;
    s->ntypes = ntypes;
    s->eps_tab = dup_d(eps_tab, kind == OR_LJ_ATTRACT_REPULSE
                                    ? (int64_t)ntypes * ntypes
                                    : ntypes);
    s->sig_tab = dup_d(sig_tab, ntypes);
    s->p3_tab = dup_d(p3_tab, ntypes);
    s->type = (int32_t *)malloc(sizeof(int32_t) * (natoms > 0 ? natoms : 1));
    memcpy(s->type, type, sizeof(int32_t) * natoms);
    s->nmembers = nmembers;
    s->member = (int32_t *)malloc(sizeof(int32_t) * (nmembers > 0 ? nmembers : 1));
    memcpy(s->member, member, sizeof(int32_t) * nmembers);
    s->diam = dup_d(diam, nmembers);
    s->skin = skin;
    s->brute = brute_rebuild;
    s->lastlocs = (double *)calloc((size_t)(nmembers > 0 ? nmembers : 1) * ndim,
                                   sizeof(double));
    s->xtmp = (double *)malloc(sizeof(double) * (natoms > 0 ? natoms : 1) * ndim);
    s->pcap = 16;
    s->first = (int32_t *)malloc(sizeof(int32_t) * s->pcap);
    s->last = (int32_t *)malloc(sizeof(int32_t) * s->pcap);
    s->npairs = 0;
    s->updatenum = 0;
    s->ignorechanged = 1; /* NeighborList ctor / add set it (trackers.hpp:194-201) */
    s->pairs = NULL;
    s->pairs_which = 0;
    return s;
}

void or_sys_destroy(or_system *s) {
    if (!s) return;
    free(s->atoms);
    free(s->eps_tab);
    free(s->sig_tab);
    free(s->p3_tab);
    free(s->type);
    free(s->member);
    free(s->diam);
    free(s->lastlocs);
    free(s->xtmp);
    free(s->first);
    free(s->last);
    free(s->pairs);
    free(s);
}

#define FOR_ATOMS(body3, body2) do {                             \
    if (s->ndim == 3) {                                          \
        atom_3d *A = (atom_3d *)s->atoms;                        \
        for (int64_t i = 0; i < s->natoms; i++)                  \
            for (int k = 0; k < 3; k++) body3;                   \
    } else {                                                     \
        atom_2d *A = (atom_2d *)s->atoms;                        \
        for (int64_t i = 0; i < s->natoms; i++)                  \
            for (int k = 0; k < 2; k++) body2;                   \
    } } while (0)

void or_sys_set_state(or_system *s, const double *x, const double *v,
                      const double *a, const double *f) {
    int d = s->ndim;
    if (x) FOR_ATOMS(A[i].x[k] = x[i * d + k], A[i].x[k] = x[i * d + k]);
    if (v) FOR_ATOMS(A[i].v[k] = v[i * d + k], A[i].v[k] = v[i * d + k]);
    if (a) FOR_ATOMS(A[i].a[k] = a[i * d + k], A[i].a[k] = a[i * d + k]);
    if (f) FOR_ATOMS(A[i].f[k] = f[i * d + k], A[i].f[k] = f[i * d + k]);
}

void or_sys_get_state(const or_system *s, double *x, double *v, double *a,
                      double *f) {
    int d = s->ndim;
    if (x) FOR_ATOMS(x[i * d + k] = A[i].x[k], x[i * d + k] = A[i].x[k]);
    if (v) FOR_ATOMS(v[i * d + k] = A[i].v[k], v[i * d + k] = A[i].v[k]);
    if (a) FOR_ATOMS(a[i * d + k] = A[i].a[k], a[i * d + k] = A[i].a[k]);
    if (f) FOR_ATOMS(f[i * d + k] = A[i].f[k], f[i * d + k] = A[i].f[k]);
}

int or_sys_update_list(or_system *s, int force) {
    return s->ndim == 3 ? sys_update_list_3d(s, force)
                        : sys_update_list_2d(s, force);
}

uint32_t or_sys_which(const or_system *s) { return s->updatenum; }
int64_t or_sys_numpairs(const or_system *s) { return s->npairs; }

int64_t or_sys_get_pairs(const or_system *s, int64_t cap, int32_t *first,
                         int32_t *last) {
    int64_t n = s->npairs < cap ? s->npairs : cap;
    memcpy(first, s->first, sizeof(int32_t) * n);
    memcpy(last, s->last, sizeof(int32_t) * n);
    return s->npairs;
}

void or_sys_set_forces(or_system *s, int and_a) {
    if (s->ndim == 3)
        sys_set_forces_3d(s, and_a);
    else
        sys_set_forces_2d(s, and_a);
}

double or_sys_potential_energy(or_system *s) {
    double U = 0;
    if (s->ndim == 3)
        sys_pair_sums_3d(s, &U, NULL, NULL);
    else
        sys_pair_sums_2d(s, &U, NULL, NULL);
    return U;
}

double or_sys_virial(or_system *s) {
    double W = 0;
    if (s->ndim == 3)
        sys_pair_sums_3d(s, NULL, &W, NULL);
    else
        sys_pair_sums_2d(s, NULL, &W, NULL);
    return W;
}

void or_sys_stress(or_system *s, double *S) {
    if (s->ndim == 3)
        sys_pair_sums_3d(s, NULL, NULL, S);
    else
        sys_pair_sums_2d(s, NULL, NULL, S);
}

double or_sys_kinetic_energy(const or_system *s, const double *originv) {
    return s->ndim == 3 ? sys_kinetic_3d(s, originv)
                        : sys_kinetic_2d(s, originv);
}

void or_sys_momentum(const or_system *s, double *p) {
    if (s->ndim == 3)
        sys_momentum_3d(s, p);
    else
        sys_momentum_2d(s, p);
}

double or_sys_mass(const or_system *s) {
    return s->ndim == 3 ? sys_mass_3d(s) : sys_mass_2d(s);
}

double or_sys_temp(const or_system *s, int minuscomv) {
    return s->ndim == 3 ? sys_temp_3d(s, minuscomv)
                        : sys_temp_2d(s, minuscomv);
}

void or_sys_initialize(or_system *s) {
    /* update_trackers; (no constraints); set_forces(true); update_trackers */
    or_sys_update_list(s, 0);
    or_sys_set_forces(s, 1);
    or_sys_update_list(s, 0);
}

void or_sys_verlet(or_system *s, double dt, int64_t nsteps) {
    if (s->ndim == 3)
        sys_verlet_3d(s, dt, nsteps);
    else
        sys_verlet_2d(s, dt, nsteps);
}

/* CollectionDamped::set_constants (src/collection.cpp:342-354). */
void or_damped_constants(double dt, double damping, double *c0, double *c1,
                         double *c2) {
    if (damping <= 0.0) {
        *c0 = 1;
        *c1 = 1;
        *c2 = .5;
        return;
    }
    double dampdt = damping * dt;
    *c0 = exp(-dampdt);
    *c1 = (-expm1(-dampdt)) / dampdt;
    *c2 = (1 - *c1) / dampdt;
}

void or_sys_damped(or_system *s, double dt, double damping, int64_t nsteps) {
    double c0, c1, c2;
    or_damped_constants(dt, damping, &c0, &c1, &c2);
    if (s->ndim == 3)
        sys_damped_3d(s, dt, c0, c1, c2, nsteps);
    else
        sys_damped_2d(s, dt, c0, c1, c2, nsteps);
}
