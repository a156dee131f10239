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
    /* NeighborList::ignorepairs (src/trackers.hpp:41-128): sorted keys
     * (hi << 32 | lo) of atom-index pairs, hi >= lo */
    int64_t nexcl, excl_cap;
    uint64_t *excl;
};

/* CollectionNLCG's members (src/collection.hpp:408-430). */
struct or_nlcg {
    int fixedL;
    double dt, seceps;
    unsigned secmax;
    double kappa, alphamax, afrac, dxmax, stepmax, kmax, P0;
    double Knew, k, vl, fl, al;
    double alpha, beta, betaused, dxsum, alphavmax, maxdV;
    unsigned sec;
};

static uint64_t excl_key(int32_t a, int32_t b) {
    uint32_t hi = (uint32_t)(a > b ? a : b), lo = (uint32_t)(a > b ? b : a);
    return ((uint64_t)hi << 32) | lo;
}

/* PairList::has_pair (src/trackers.hpp:66-71): the unordered pair is stored. */
static int or_excluded(int64_t n, const uint64_t *keys, int32_t a, int32_t b) {
    if (n == 0) return 0;
    uint64_t k = excl_key(a, b);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (keys[mid] == k) return 1;
        if (keys[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return 0;
}

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

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Sorted, de-duplicated exclusion keys of nexcl (ea[k], eb[k]) pairs. */
static int64_t excl_keys(int64_t nexcl, const int32_t *ea, const int32_t *eb,
                         uint64_t **out) {
    uint64_t *k = (uint64_t *)malloc(sizeof(uint64_t) * (nexcl > 0 ? nexcl : 1));
    for (int64_t q = 0; q < nexcl; q++) k[q] = excl_key(ea[q], eb[q]);
    qsort(k, (size_t)nexcl, sizeof(uint64_t), cmp_u64);
    int64_t m = 0;
    for (int64_t q = 0; q < nexcl; q++)
        if (m == 0 || k[q] != k[m - 1]) k[m++] = k[q];
    *out = k;
    return m;
}

int64_t or_nlist_brute(int ndim, int64_t M, const int32_t *member,
                       const double *diam, const double *x, const double *L,
                       double skin, int64_t cap, int32_t *first,
                       int32_t *last, int64_t nexcl, const int32_t *ea,
                       const int32_t *eb) {
    uint64_t *k;
    int64_t m = excl_keys(nexcl, ea, eb, &k);
    int64_t r = ndim == 3 ? nlist_brute_3d(M, member, diam, x, L, skin, cap,
                                           first, last, m, k)
                          : nlist_brute_2d(M, member, diam, x, L, skin, cap,
                                           first, last, m, k);
    free(k);
    return r;
}

int64_t or_nlist_cells(int ndim, int64_t M, const int32_t *member,
                       const double *diam, const double *x, const double *L,
                       double skin, int64_t cap, int32_t *first,
                       int32_t *last, int64_t nexcl, const int32_t *ea,
                       const int32_t *eb) {
    uint64_t *k;
    int64_t m = excl_keys(nexcl, ea, eb, &k);
    int64_t r = ndim == 3 ? nlist_cells_3d(M, member, diam, x, L, skin, cap,
                                           first, last, m, k)
                          : nlist_cells_2d(M, member, diam, x, L, skin, cap,
                                           first, last, m, k);
    free(k);
    return r;
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
    const int eps_rows = kind == OR_LJ_ATTRACT_REPULSE || kind == OR_LJ_CUT_II ||
                         kind == OR_REPULSION_II;
    const int sig_rows = kind == OR_LJ_CUT_II || kind == OR_REPULSION_II;
    s->eps_tab = dup_d(eps_tab, eps_rows ? (int64_t)ntypes * ntypes : ntypes);
    s->sig_tab = dup_d(sig_tab, sig_rows ? (int64_t)ntypes * ntypes : ntypes);
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
    free(s->excl);
    free(s);
}

/* NeighborList::ignore(a, b) (src/trackers.hpp:190-193): PairList::add_pair
 * (:76-83) stores the unordered pair once; the next update rebuilds. */
void or_sys_ignore(or_system *s, int32_t a, int32_t b) {
    uint64_t k = excl_key(a, b);
    int64_t pos = 0;
    while (pos < s->nexcl && s->excl[pos] < k) pos++;
    if (pos < s->nexcl && s->excl[pos] == k) {
        s->ignorechanged = 1;
        return;
    }
    if (s->nexcl == s->excl_cap) {
        s->excl_cap = s->excl_cap ? 2 * s->excl_cap : 16;
        s->excl = (uint64_t *)realloc(s->excl, sizeof(uint64_t) * s->excl_cap);
    }
    memmove(s->excl + pos + 1, s->excl + pos, sizeof(uint64_t) * (s->nexcl - pos));
    s->excl[pos] = k;
    s->nexcl++;
    s->ignorechanged = 1;
}

/* NeighborList::ignore_size (src/trackers.hpp:204, PairList::size :117-123). */
int64_t or_sys_ignore_size(const or_system *s) { return s->nexcl; }

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

/* ------------------------------------------------------------------ NLCG */
or_nlcg *or_nlcg_create(or_system *s, double dt, double P0, double kappa,
                        double kmax, int secmax, double seceps, int fixedL) {
    /* Collection's constructor runs initialize() (src/collection.cpp:3-19)
     * with Collection::set_forces; the NLCG members start as in
     * src/collection.cpp:500-523 */
    or_sys_initialize(s);
    or_nlcg *c = (or_nlcg *)calloc(1, sizeof(or_nlcg));
    c->fixedL = fixedL;
    c->dt = dt;
    c->P0 = P0;
    c->kappa = kappa;
    c->kmax = kmax;
    c->secmax = (unsigned)secmax;
    c->seceps = seceps;
    c->alphamax = 2.0;
    c->afrac = 0;
    c->dxmax = 100;
    c->stepmax = 1e-3;
    return c;
}

void or_nlcg_destroy(or_nlcg *c) { free(c); }

void or_nlcg_set_limits(or_nlcg *c, double alphamax, double afrac, double dxmax,
                        double stepmax, double maxdV) {
    c->alphamax = alphamax;
    c->afrac = afrac;
    c->dxmax = dxmax;
    c->stepmax = stepmax;
    c->maxdV = maxdV;
}

void or_nlcg_set_forces(or_system *s, or_nlcg *c, int cna, int setV) {
    if (s->ndim == 3)
        nlcg_set_forces_3d(s, c, cna, setV);
    else
        nlcg_set_forces_2d(s, c, cna, setV);
}

void or_nlcg_reset(or_system *s, or_nlcg *c) {
    /* reset (src/collection.cpp:528-535) */
    c->k = 0;
    or_nlcg_set_forces(s, c, 1, 1);
    c->vl = c->al;
}

void or_nlcg_timestep(or_system *s, or_nlcg *c, int64_t nsteps) {
    for (int64_t q = 0; q < nsteps; q++) {
        if (s->ndim == 3)
            nlcg_timestep_3d(s, c);
        else
            nlcg_timestep_2d(s, c);
    }
}

void or_nlcg_descend(or_system *s, or_nlcg *c) {
    if (s->ndim == 3)
        nlcg_descend_3d(s, c);
    else
        nlcg_descend_2d(s, c);
}

void or_nlcg_observables(or_system *s, or_nlcg *c, double *H, double *P, double *K) {
    double U = 0, W = 0, V;
    if (s->ndim == 3) {
        sys_pair_sums_3d(s, &U, &W, NULL);
        V = box_V_3d(s);
        *K = nlcg_kinetic_3d(s, c);
    } else {
        sys_pair_sums_2d(s, &U, &W, NULL);
        V = box_V_2d(s);
        *K = nlcg_kinetic_2d(s, c);
    }
    *H = U + c->P0 * V;
    *P = W / V / (double)s->ndim;
}

void or_nlcg_info(const or_nlcg *c, double *out) {
    out[0] = c->alpha;
    out[1] = c->beta;
    out[2] = c->betaused;
    out[3] = c->dxsum;
    out[4] = c->alphavmax;
    out[5] = c->sec;
    out[6] = c->k;
    out[7] = c->vl;
    out[8] = c->fl;
    out[9] = c->al;
    out[10] = c->Knew;
    out[11] = c->maxdV;
}

void or_sys_box(const or_system *s, double *L) {
    for (int k = 0; k < s->ndim; k++) L[k] = s->L[k];
}
