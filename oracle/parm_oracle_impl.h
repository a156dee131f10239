/*
 * parm_oracle_impl.h -- dimension-specialised body of the CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see parm_oracle.h).  Included twice by
 * parm_oracle.c with NDIM = 2 and NDIM = 3; every symbol gets the suffix
 * _2d / _3d through SUF().
 */

/* One atom in the reference's AoS layout: struct Atom {Vec x, v, a, f; flt m;}
 * (src/box.hpp:234-249). */
typedef struct {
    double x[NDIM], v[NDIM], a[NDIM], f[NDIM];
    double m;
} SUF(atom);

/* Materialised LJAttractRepulsePair / RepulsionPair (src/interaction.hpp:
 * 1251-1299, 1528-1590), built by update_pairs (src/interaction.hpp:
 * 2101-2115) whenever the list's which() changes. */
typedef struct {
    int32_t i, j;          /* atom1 (later-added member), atom2 */
    double eps, sig, p3, p4; /* LJ: cut, cut_energy ; Repulsion: exponent, - */
} SUF(pairP);

/* r = diff(r1, r2) = vec_mod(r1 - r2, L)  (src/box.hpp:68-78, 103). */
static inline void SUF(diff)(const double *r1, const double *r2,
                             const double *L, double *r) {
    for (int k = 0; k < NDIM; k++) r[k] = remainder(r1[k] - r2[k], L[k]);
}

/* Vec::squaredNorm in the Eigen vectorised order (x*x + y*y) + z*z. */
static inline double SUF(sqnorm)(const double *r) {
#if NDIM == 3
    return (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
#else
    return r[0] * r[0] + r[1] * r[1];
#endif
}

static inline double SUF(dot)(const double *r, const double *f) {
#if NDIM == 3
    return (r[0] * f[0] + r[1] * f[1]) + r[2] * f[2];
#else
    return r[0] * f[0] + r[1] * f[1];
#endif
}

/* Pair parameters exactly as the pair constructors compute them.
 * LJAttractRepulsePair(a1, a2): eps = a1.epsilons[a2.indx],
 *   sig = (s1 + s2) / 2.0, cut = max(c1, c2); eps <= 0 -> cut 1, cut_energy 0,
 *   eps = |eps|; else cut_energy = eps * (1 - cut^-6)^2
 *   (src/interaction.hpp:1256-1269).
 * RepulsionPair(a1, a2): eps = sqrt(e1 e2), sig = (s1 + s2) / 2.0,
 *   exponent = (n1 + n2) / 2.0  (src/interaction.hpp:1532-1537); with
 *   IEpsISigExpAtom eps = a1.epsilons[a2.indx], sig = a1.sigmas[a2.indx]
 *   (:1537-1542, get_epsilon / get_sigma :1500-1513).
 * LennardJonesCutPair(a1, a2): EpsSigCutAtom eps = sqrt(e1 e2),
 *   sig = (s1 + s2) / 2; IEpsISigCutAtom eps / sig from the rows; cut =
 *   max(c1, c2); LennardJonesCut's cut_energy = eps ((1 - 1/cut^6)^2 - 1)
 *   (src/interaction.hpp:970-984, 249-253).
 * LJRepulsePair(a1, a2): eps = sqrt(e1 e2), sig = (s1 + s2) / 2
 *   (src/interaction.hpp:875-883). */
static void SUF(pair_params)(int kind, int ntypes, const double *eps_tab,
                             const double *sig_tab, const double *p3_tab,
                             int t1, int t2, SUF(pairP) *p) {
    const int64_t ix = (int64_t)t1 * ntypes + t2;
    if (kind == OR_LJ_CUT || kind == OR_LJ_CUT_II) {
        double c1 = p3_tab[t1], c2 = p3_tab[t2];
        double cut = (c1 < c2) ? c2 : c1; /* std::max */
        if (kind == OR_LJ_CUT) {
            p->eps = sqrt(eps_tab[t1] * eps_tab[t2]);
            p->sig = (sig_tab[t1] + sig_tab[t2]) / 2;
        } else {
            p->eps = eps_tab[ix];
            p->sig = sig_tab[ix];
        }
        double rsix = pow(cut, 6);
        double mid = (1 - 1 / rsix);
        p->p3 = cut;
        p->p4 = p->eps * (mid * mid - 1);
        return;
    }
    if (kind == OR_LJ_REPULSE) {
        p->eps = sqrt(eps_tab[t1] * eps_tab[t2]);
        p->sig = (sig_tab[t1] + sig_tab[t2]) / 2;
        p->p3 = 1;
        p->p4 = 0;
        return;
    }
    if (kind == OR_REPULSION_II) {
        p->eps = eps_tab[ix];
        p->sig = sig_tab[ix];
        p->p3 = (p3_tab[t1] + p3_tab[t2]) / 2.0;
        p->p4 = 0;
        return;
    }
    if (kind == OR_LJ_ATTRACT_REPULSE) {
        double eps = eps_tab[(int64_t)t1 * ntypes + t2];
        double sig = (sig_tab[t1] + sig_tab[t2]) / 2.0;
        double c1 = p3_tab[t1], c2 = p3_tab[t2];
        double cut = (c1 < c2) ? c2 : c1; /* std::max */
        double cut_energy;
        if (eps <= 0) {
            cut = 1;
            cut_energy = 0;
            eps = fabs(eps);
        } else {
            double mid = (1 - pow(cut, -6));
            cut_energy = eps * (mid * mid);
        }
        p->eps = eps;
        p->sig = sig;
        p->p3 = cut;
        p->p4 = cut_energy;
    } else {
        p->eps = sqrt(eps_tab[t1] * eps_tab[t2]);
        p->sig = (sig_tab[t1] + sig_tab[t2]) / 2.0;
        p->p3 = (p3_tab[t1] + p3_tab[t2]) / 2.0;
        p->p4 = 0;
    }
}

/* Force on atom1 (into f) for displacement rij = diff(x1, x2); returns the
 * pair energy.  LJAttractRepulsePair::forces / energy
 * (src/interaction.hpp:1270-1298), RepulsionPair::forces / energy
 * (src/interaction.hpp:1544-1556), LennardJonesCut::forces / energy
 * (src/interaction.hpp:254-272) and LJRepulsive::forces / energy
 * (src/interaction.hpp:126-147). */
static inline double SUF(pair_eval)(int kind, const SUF(pairP) *p,
                                    const double *rij, double *f) {
    for (int k = 0; k < NDIM; k++) f[k] = 0;
    double dsq = SUF(sqnorm)(rij);
    if (kind == OR_LJ_CUT || kind == OR_LJ_CUT_II || kind == OR_LJ_REPULSE) {
        /* LennardJonesCut (cut = p3) and LJRepulsive (cut 1, no shift) share
         * their arithmetic: rsix = (r^2/sig^2)^3, 12 eps / rsix (1/rsix - 1) */
        double rsq = dsq / (p->sig * p->sig);
        if (rsq > (p->p3 * p->p3)) return 0;
        double rsix = rsq * rsq * rsq;
        double fmagTimesR = 12 * p->eps / rsix * (1 / rsix - 1);
        double s = fmagTimesR / dsq;
        for (int k = 0; k < NDIM; k++) f[k] = rij[k] * s;
        double mid = (1 - 1 / rsix);
        if (kind == OR_LJ_REPULSE) return p->eps * (mid * mid);
        return p->eps * (mid * mid - 1) - p->p4;
    }
    if (kind == OR_LJ_ATTRACT_REPULSE) {
        double rsq = dsq / (p->sig * p->sig);
        if (rsq > p->p3 * p->p3) return 0;
        double rsix = pow(rsq, -3); /* sigma^6 / r^6 */
        if (p->eps != 0) {
            double fmagTimesR = 12 * p->eps * rsix * (rsix - 1);
            double s = fmagTimesR / dsq;
            for (int k = 0; k < NDIM; k++) f[k] = rij[k] * s;
        }
        double mid = (1 - rsix);
        return p->eps * (mid * mid) - p->p4;
    } else {
        if (dsq > p->sig * p->sig) return 0;
        double R = sqrt(dsq);
        double s = p->eps * pow(1.0 - (R / p->sig), p->p3 - 1) / p->sig / R;
        for (int k = 0; k < NDIM; k++) f[k] = rij[k] * s;
        return p->eps * pow(1.0 - (R / p->sig), p->p3) / p->p3;
    }
}

/* NeighborList pair predicate (src/trackers.cpp:63-67). */
static inline int SUF(in_list)(const double *xi, const double *xj,
                               const double *L, double di, double dj,
                               double skin) {
    double r[NDIM];
    SUF(diff)(xi, xj, L, r);
    double diam = (di + dj) / 2;
    return sqrt(SUF(sqnorm)(r)) < (diam + skin);
}

static int64_t SUF(nlist_brute)(int64_t M, const int32_t *member,
                                const double *diam, const double *x,
                                const double *L, double skin, int64_t cap,
                                int32_t *first, int32_t *last, int64_t nexcl,
                                const uint64_t *excl) {
    int64_t np = 0;
    for (int64_t i = 0; i < M; i++) {
        const double *xi = x + (int64_t)member[i] * NDIM;
        for (int64_t j = 0; j < i; j++) {
            const double *xj = x + (int64_t)member[j] * NDIM;
            if (or_excluded(nexcl, excl, member[i], member[j])) continue;
            if (SUF(in_list)(xi, xj, L, diam[i], diam[j], skin)) {
                if (np < cap) {
                    first[np] = member[i];
                    last[np] = member[j];
                }
                np++;
            }
        }
    }
    return np;
}

static int SUF(cmp_i32)(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Periodic cell grid with cell edge >= the largest list radius: every pair
 * within the list radius lies in the same or an adjacent cell. */
static int64_t SUF(nlist_cells)(int64_t M, const int32_t *member,
                                const double *diam, const double *x,
                                const double *L, double skin, int64_t cap,
                                int32_t *first, int32_t *last, int64_t nexcl,
                                const uint64_t *excl) {
    if (M <= 0) return 0;
    double dmax = 0;
    for (int64_t i = 0; i < M; i++)
        if (diam[i] > dmax) dmax = diam[i];
    double rmax = (dmax + skin) * (1 + 1e-9);
    int nc[NDIM];
    int64_t ncell = 1;
    for (int k = 0; k < NDIM; k++) {
        double q = floor(L[k] / rmax);
        nc[k] = (q >= 3) ? (q > 1024 ? 1024 : (int)q) : 1;
        ncell *= nc[k];
    }
    int32_t *cell = (int32_t *)malloc(sizeof(int32_t) * M);
    int64_t *head = (int64_t *)malloc(sizeof(int64_t) * (ncell + 1));
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * M);
    for (int64_t c = 0; c <= ncell; c++) head[c] = 0;
    for (int64_t i = 0; i < M; i++) {
        const double *xi = x + (int64_t)member[i] * NDIM;
        int64_t c = 0;
        for (int k = NDIM - 1; k >= 0; k--) {
            double s = xi[k] / L[k];
            s -= floor(s);
            int ck = (int)(s * nc[k]);
            if (ck >= nc[k]) ck = nc[k] - 1;
            if (ck < 0) ck = 0;
            c = c * nc[k] + ck;
        }
        cell[i] = (int32_t)c;
        head[c + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) head[c + 1] += head[c];
    {
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ncell);
        for (int64_t c = 0; c < ncell; c++) fill[c] = head[c];
        for (int64_t i = 0; i < M; i++) order[fill[cell[i]]++] = (int32_t)i;
        free(fill);
    }
    int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (M > 0 ? M : 1));
    int64_t np = 0;
    for (int64_t i = 0; i < M; i++) {
        const double *xi = x + (int64_t)member[i] * NDIM;
        int ci[NDIM];
        int64_t c = cell[i];
        for (int k = 0; k < NDIM; k++) {
            ci[k] = (int)(c % nc[k]);
            c /= nc[k];
        }
        int64_t ncand = 0;
        int lo[NDIM], hi[NDIM], o[NDIM];
        for (int k = 0; k < NDIM; k++) {
            lo[k] = nc[k] >= 3 ? -1 : 0;
            hi[k] = nc[k] >= 3 ? 1 : 0;
            o[k] = lo[k];
        }
        for (;;) {
            int64_t cc = 0;
            for (int k = NDIM - 1; k >= 0; k--) {
                int q = ci[k] + o[k];
                if (q < 0) q += nc[k];
                if (q >= nc[k]) q -= nc[k];
                cc = cc * nc[k] + q;
            }
            for (int64_t t = head[cc]; t < head[cc + 1]; t++) {
                int32_t j = order[t];
                if (j >= i) continue;
                if (or_excluded(nexcl, excl, member[i], member[j])) continue;
                const double *xj = x + (int64_t)member[j] * NDIM;
                if (SUF(in_list)(xi, xj, L, diam[i], diam[j], skin))
                    cand[ncand++] = j;
            }
            int k = 0;
            while (k < NDIM) {
                if (++o[k] <= hi[k]) break;
                o[k] = lo[k];
                k++;
            }
            if (k == NDIM) break;
        }
        qsort(cand, (size_t)ncand, sizeof(int32_t), SUF(cmp_i32));
        for (int64_t t = 0; t < ncand; t++) {
            if (np < cap) {
                first[np] = member[i];
                last[np] = member[cand[t]];
            }
            np++;
        }
    }
    free(cand);
    free(order);
    free(head);
    free(cell);
    return np;
}

static int SUF(skin_check)(int64_t M, const int32_t *member, const double *x,
                           const double *lastlocs, double skin) {
    double bigdist = 0, biggestdist = 0;
    for (int64_t i = 0; i < M; i++) {
        const double *xi = x + (int64_t)member[i] * NDIM;
        double r[NDIM];
        for (int k = 0; k < NDIM; k++) r[k] = xi[k] - lastlocs[i * NDIM + k];
        double curdist = sqrt(SUF(sqnorm)(r));
        if (curdist > biggestdist) {
            bigdist = biggestdist;
            biggestdist = curdist;
        } else if (curdist > bigdist) {
            bigdist = curdist;
        } else
            continue;
        if (bigdist + biggestdist >= skin) return 1;
    }
    return 0;
}

static void SUF(forces)(int64_t natoms, const double *x, const double *L,
                        int kind, int ntypes, const double *eps_tab,
                        const double *sig_tab, const double *p3_tab,
                        const int32_t *type, int64_t npairs,
                        const int32_t *first, const int32_t *last, double *f,
                        double *fabs_sum, double *U, double *W, double *S,
                        uint64_t *contacts, uint64_t *overlaps) {
    (void)natoms;
    double E = 0, P = 0, St[NDIM * NDIM];
    uint64_t nc = 0, no = 0;
    for (int k = 0; k < NDIM * NDIM; k++) St[k] = 0;
    for (int64_t p = 0; p < npairs; p++) {
        int32_t i = first[p], j = last[p];
        SUF(pairP) pp;
        SUF(pair_params)(kind, ntypes, eps_tab, sig_tab, p3_tab, type[i],
                         type[j], &pp);
        double r[NDIM], fp[NDIM];
        SUF(diff)(x + (int64_t)i * NDIM, x + (int64_t)j * NDIM, L, r);
        double e = SUF(pair_eval)(kind, &pp, r, fp);
        double fa = sqrt(SUF(sqnorm)(fp));
        for (int k = 0; k < NDIM; k++) {
            f[(int64_t)i * NDIM + k] += fp[k];
            f[(int64_t)j * NDIM + k] -= fp[k];
        }
        if (fabs_sum) {
            fabs_sum[i] += fa;
            fabs_sum[j] += fa;
        }
        E += e;
        if (e != 0.0) nc++;
        if (e > 0.0) no++;
        P += SUF(dot)(r, fp);
        for (int a = 0; a < NDIM; a++)
            for (int b = 0; b < NDIM; b++) St[a * NDIM + b] += r[a] * fp[b];
    }
    if (U) *U = E;
    if (W) *W = P;
    if (S)
        for (int k = 0; k < NDIM * NDIM; k++) S[k] = St[k];
    if (contacts) *contacts = nc;
    if (overlaps) *overlaps = no;
}

/* ------------------------------------------------------------------ system */

static int SUF(is_fixed)(double m) { return m <= 0 || isinf(m); }

static void SUF(sys_rebuild)(or_system *s) {
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    double *xs = s->xtmp;
    for (int64_t i = 0; i < s->natoms; i++)
        for (int k = 0; k < NDIM; k++) xs[i * NDIM + k] = A[i].x[k];
    /* the reference stores lastlocs[i] = x_i inside the rebuild loop */
    for (int64_t i = 0; i < s->nmembers; i++)
        for (int k = 0; k < NDIM; k++)
            s->lastlocs[i * NDIM + k] = A[s->member[i]].x[k];
    int64_t np;
    for (;;) {
        if (s->brute)
            np = SUF(nlist_brute)(s->nmembers, s->member, s->diam, xs, s->L,
                                  s->skin, s->pcap, s->first, s->last,
                                  s->nexcl, s->excl);
        else
            np = SUF(nlist_cells)(s->nmembers, s->member, s->diam, xs, s->L,
                                  s->skin, s->pcap, s->first, s->last,
                                  s->nexcl, s->excl);
        if (np <= s->pcap) break;
        s->pcap = np + np / 4 + 16;
        s->first = (int32_t *)realloc(s->first, sizeof(int32_t) * s->pcap);
        s->last = (int32_t *)realloc(s->last, sizeof(int32_t) * s->pcap);
    }
    s->npairs = np;
    s->updatenum++;
}

static int SUF(sys_update_list)(or_system *s, int force) {
    if (!force && !s->ignorechanged) {
        double *xs = s->xtmp;
        SUF(atom) *A = (SUF(atom) *)s->atoms;
        for (int64_t i = 0; i < s->natoms; i++)
            for (int k = 0; k < NDIM; k++) xs[i * NDIM + k] = A[i].x[k];
        if (!SUF(skin_check)(s->nmembers, s->member, xs, s->lastlocs, s->skin))
            return 0;
    }
    s->ignorechanged = 0;
    SUF(sys_rebuild)(s);
    return 1;
}

/* NListed::update_pairs (src/interaction.hpp:2101-2115). */
static void SUF(sys_update_pairs)(or_system *s) {
    if (s->pairs_which == s->updatenum && s->pairs) return;
    s->pairs_which = s->updatenum;
    free(s->pairs);
    SUF(pairP) *P = (SUF(pairP) *)malloc(sizeof(SUF(pairP)) *
                                         (s->npairs > 0 ? s->npairs : 1));
    for (int64_t p = 0; p < s->npairs; p++) {
        SUF(pair_params)(s->kind, s->ntypes, s->eps_tab, s->sig_tab,
                         s->p3_tab, s->type[s->first[p]],
                         s->type[s->last[p]], &P[p]);
        P[p].i = s->first[p];
        P[p].j = s->last[p];
    }
    s->pairs = P;
}

/* Collection::set_forces (src/collection.cpp:159-179) with NListed::set_forces
 * (src/interaction.hpp:2165-2175). */
static void SUF(sys_set_forces)(or_system *s, int and_a) {
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    for (int64_t i = 0; i < s->natoms; i++)
        for (int k = 0; k < NDIM; k++) A[i].f[k] = 0;
    SUF(sys_update_pairs)(s);
    const SUF(pairP) *P = (const SUF(pairP) *)s->pairs;
    for (int64_t p = 0; p < s->npairs; p++) {
        double r[NDIM], fp[NDIM];
        SUF(atom) *a1 = &A[P[p].i], *a2 = &A[P[p].j];
        SUF(diff)(a1->x, a2->x, s->L, r);
        SUF(pair_eval)(s->kind, &P[p], r, fp);
        for (int k = 0; k < NDIM; k++) {
            a1->f[k] += fp[k];
            a2->f[k] -= fp[k];
        }
    }
    if (!and_a) return;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) {
            for (int k = 0; k < NDIM; k++) A[i].a[k] = 0;
            continue;
        }
        for (int k = 0; k < NDIM; k++) A[i].a[k] = A[i].f[k] / A[i].m;
    }
}

static void SUF(sys_pair_sums)(or_system *s, double *U, double *W, double *S) {
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    SUF(sys_update_pairs)(s);
    const SUF(pairP) *P = (const SUF(pairP) *)s->pairs;
    double E = 0, Pr = 0, St[NDIM * NDIM];
    for (int k = 0; k < NDIM * NDIM; k++) St[k] = 0;
    for (int64_t p = 0; p < s->npairs; p++) {
        double r[NDIM], fp[NDIM];
        SUF(diff)(A[P[p].i].x, A[P[p].j].x, s->L, r);
        E += SUF(pair_eval)(s->kind, &P[p], r, fp);
        Pr += SUF(dot)(r, fp);
        for (int a = 0; a < NDIM; a++)
            for (int b = 0; b < NDIM; b++) St[a * NDIM + b] += r[a] * fp[b];
    }
    if (U) *U = E;
    if (W) *W = Pr;
    if (S)
        for (int k = 0; k < NDIM * NDIM; k++) S[k] = St[k];
}

static double SUF(sys_kinetic)(const or_system *s, const double *ov) {
    const SUF(atom) *A = (const SUF(atom) *)s->atoms;
    double totE = 0;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (A[i].m == 0 || isinf(A[i].m)) continue;
        double c[NDIM];
        for (int k = 0; k < NDIM; k++) c[k] = A[i].v[k] - (ov ? ov[k] : 0.0);
        totE += A[i].m / 2 * SUF(dot)(c, c);
    }
    return totE;
}

static void SUF(sys_momentum)(const or_system *s, double *p) {
    const SUF(atom) *A = (const SUF(atom) *)s->atoms;
    for (int k = 0; k < NDIM; k++) p[k] = 0;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) continue;
        for (int k = 0; k < NDIM; k++) p[k] += A[i].v[k] * A[i].m;
    }
}

static double SUF(sys_mass)(const or_system *s) {
    const SUF(atom) *A = (const SUF(atom) *)s->atoms;
    double m = 0;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) continue;
        m += A[i].m;
    }
    return m;
}

/* Collection::temp (src/collection.cpp:116-142). */
static double SUF(sys_temp)(const or_system *s, int minuscomv) {
    const SUF(atom) *A = (const SUF(atom) *)s->atoms;
    double v[NDIM];
    for (int k = 0; k < NDIM; k++) v[k] = 0;
    if (minuscomv) {
        double p[NDIM], M = SUF(sys_mass)(s);
        SUF(sys_momentum)(s, p);
        for (int k = 0; k < NDIM; k++) v[k] = p[k] / M;
    }
    int ndof = 0;
    for (int64_t i = 0; i < s->natoms; i++)
        if (!SUF(is_fixed)(A[i].m)) ndof += NDIM;
    if (minuscomv) ndof -= NDIM;
    return SUF(sys_kinetic)(s, v) * 2 / ndof;
}

/* CollectionVerlet::timestep (src/collection.cpp:442-469). */
static void SUF(sys_verlet)(or_system *s, double dt, int64_t nsteps) {
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    for (int64_t st = 0; st < nsteps; st++) {
        for (int64_t i = 0; i < s->natoms; i++) {
            if (SUF(is_fixed)(A[i].m)) {
                for (int k = 0; k < NDIM; k++) A[i].v[k] = 0;
                continue;
            }
            for (int k = 0; k < NDIM; k++) {
                A[i].x[k] += A[i].v[k] * dt + A[i].a[k] * (dt * dt / 2);
                A[i].v[k] += A[i].a[k] * (dt / 2);
            }
        }
        SUF(sys_set_forces)(s, 0);
        for (int64_t i = 0; i < s->natoms; i++) {
            if (SUF(is_fixed)(A[i].m)) {
                for (int k = 0; k < NDIM; k++) A[i].a[k] = 0;
                continue;
            }
            for (int k = 0; k < NDIM; k++) {
                A[i].a[k] = A[i].f[k] / A[i].m;
                A[i].v[k] += A[i].a[k] * (dt / 2);
            }
        }
        SUF(sys_update_list)(s, 0);
    }
}

/* CollectionDamped::timestep (src/collection.cpp:356-381). */
static void SUF(sys_damped)(or_system *s, double dt, double c0, double c1,
                            double c2, int64_t nsteps) {
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    for (int64_t st = 0; st < nsteps; st++) {
        for (int64_t i = 0; i < s->natoms; i++) {
            if (SUF(is_fixed)(A[i].m)) {
                for (int k = 0; k < NDIM; k++) A[i].v[k] = 0;
                continue;
            }
            for (int k = 0; k < NDIM; k++) {
                A[i].x[k] += A[i].v[k] * (c1 * dt) + A[i].a[k] * (c2 * dt * dt);
                A[i].v[k] = A[i].v[k] * c0 + A[i].a[k] * (dt * (c1 - c2));
            }
        }
        SUF(sys_set_forces)(s, 1);
        for (int64_t i = 0; i < s->natoms; i++) {
            if (SUF(is_fixed)(A[i].m)) continue;
            for (int k = 0; k < NDIM; k++) A[i].v[k] += A[i].a[k] * (dt * c2);
        }
        SUF(sys_update_list)(s, 0);
    }
}

/* ------------------------------------------------------------------ NLCG */
/* CollectionNLCG / CollectionNLCGFixedL (src/collection.cpp:494-870), with
 * the system's one NListed interaction and its NeighborList as the only
 * tracker.  Atom loops in atom order, sums in that order. */

/* Collection::set_forces_get_pressure(false) (src/collection.cpp:181-208)
 * with NListed::set_forces_get_pressure (src/interaction.hpp:2231-2245). */
static double SUF(sys_forces_get_pressure)(or_system *s) {
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    for (int64_t i = 0; i < s->natoms; i++)
        for (int k = 0; k < NDIM; k++) A[i].f[k] = 0;
    SUF(sys_update_pairs)(s);
    const SUF(pairP) *P = (const SUF(pairP) *)s->pairs;
    double p = 0;
    for (int64_t q = 0; q < s->npairs; q++) {
        double r[NDIM], fp[NDIM];
        SUF(atom) *a1 = &A[P[q].i], *a2 = &A[P[q].j];
        SUF(diff)(a1->x, a2->x, s->L, r);
        SUF(pair_eval)(s->kind, &P[q], r, fp);
        for (int k = 0; k < NDIM; k++) {
            a1->f[k] += fp[k];
            a2->f[k] -= fp[k];
        }
        p += SUF(dot)(r, fp);
    }
    for (int64_t i = 0; i < s->natoms; i++)
        if (SUF(is_fixed)(A[i].m))
            for (int k = 0; k < NDIM; k++) A[i].a[k] = 0;
    return p;
}

static double SUF(box_V)(const or_system *s) {
    double V = 1;
    for (int k = 0; k < NDIM; k++) V *= s->L[k];
    return V;
}

/* get_length_squared (src/collection.cpp:627-635) */
static double SUF(nlcg_l2)(const or_system *s) {
#if NDIM == 3
    return pow(SUF(box_V)(s), 2.0 / 3.0);
#else
    return SUF(box_V)(s);
#endif
}

/* fdotf / fdota / fdotv / vdotv (src/collection.cpp:637-683):
 * which = 0 f.f, 1 f.a, 2 f.v, 3 v.v over atoms that are not fixed. */
static double SUF(nlcg_dot)(const or_system *s, const or_nlcg *c, int which) {
    const SUF(atom) *A = (const SUF(atom) *)s->atoms;
    double r = 0;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) continue;
        const double *p = which == 3 ? A[i].v : A[i].f;
        const double *q = which == 0 ? A[i].f : which == 1 ? A[i].a : A[i].v;
        r += SUF(dot)(p, q);
    }
    double sc = which == 0 ? c->fl * c->fl
                : which == 1 ? c->fl * c->al
                : which == 2 ? c->fl * c->vl
                             : c->vl * c->vl;
    return r / SUF(nlcg_l2)(s) + sc;
}

/* stepx (src/collection.cpp:608-625; FixedL :855-870) */
static void SUF(nlcg_stepx)(or_system *s, or_nlcg *c, double dx) {
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    double Lfac = c->fixedL ? exp(dx * c->vl / c->kappa)
                            : exp(dx * c->vl / (c->kappa * NDIM));
    for (int64_t i = 0; i < s->natoms; i++) {
        if (c->fixedL)
            A[i].x[0] *= Lfac;
        else
            for (int k = 0; k < NDIM; k++) A[i].x[k] *= Lfac;
        for (int k = 0; k < NDIM; k++) A[i].x[k] += A[i].v[k] * dx;
    }
    if (c->fixedL)
        s->L[0] *= Lfac;
    else
        for (int k = 0; k < NDIM; k++) s->L[k] *= Lfac;
}

/* CollectionNLCG::set_forces(constraints_and_a, setV) (:537-569) */
static void SUF(nlcg_set_forces)(or_system *s, or_nlcg *c, int cna, int setV) {
    double V = SUF(box_V)(s);
    if (setV) {
        double interacP = SUF(sys_forces_get_pressure)(s);
        c->fl = ((interacP / NDIM) - (c->P0 * V)) / c->kappa;
        if (cna) {
            c->al = c->fl;
            c->vl = c->fl;
        }
    } else {
        SUF(sys_set_forces)(s, 0);
    }
    if (cna) {
        SUF(atom) *A = (SUF(atom) *)s->atoms;
        for (int64_t i = 0; i < s->natoms; i++) {
            if (SUF(is_fixed)(A[i].m)) {
                for (int k = 0; k < NDIM; k++) A[i].v[k] = A[i].a[k] = 0;
                continue;
            }
            for (int k = 0; k < NDIM; k++) A[i].a[k] = A[i].v[k] = A[i].f[k];
        }
        c->Knew = SUF(nlcg_dot)(s, c, 1);
    }
}

/* CollectionNLCG::timestep (src/collection.cpp:685-835) */
static void SUF(nlcg_timestep)(or_system *s, or_nlcg *c) {
    SUF(nlcg_stepx)(s, c, c->dt);
    SUF(nlcg_set_forces)(s, c, 0, 1);
    double eta0 = -SUF(nlcg_dot)(s, c, 2);
    double eta;
    SUF(nlcg_stepx)(s, c, -c->dt);
    SUF(sys_update_list)(s, 0);
    c->alpha = -c->dt;
    SUF(nlcg_set_forces)(s, c, 0, 1);
    c->dxsum = 0;
    double vdv = SUF(nlcg_dot)(s, c, 3);
    for (c->sec = 0; c->sec < c->secmax; c->sec++) {
        eta = -SUF(nlcg_dot)(s, c, 2);
        double alphafac = -eta / fabs(eta0 - eta);
        if (fabs(eta0 - eta) <= 1e-12 * fabs(eta)) {
            alphafac = c->alphamax > 0 ? c->alphamax : 1.1;
            c->sec = c->sec > 0 ? c->sec * 2 - 1 : 1;
        }
        if (c->alphamax > 0 && alphafac > c->alphamax) alphafac = c->alphamax;
        if (c->alphamax > 0 && alphafac < -c->alphamax) alphafac = -c->alphamax;
        c->alpha = fabs(c->alpha) * alphafac;
        double newdxsum = fabs(c->dxsum + c->alpha);
        if (c->dxmax > 0 && newdxsum > c->dxmax) {
            c->k = 0;
            break;
        }
        double dVoverV = expm1(fabs(c->dxsum + c->alpha) * c->vl / (c->kappa * NDIM));
        if (c->maxdV > 0 && dVoverV > c->maxdV) {
            c->k = 0;
            break;
        }
        c->dxsum += c->alpha;
        SUF(nlcg_stepx)(s, c, c->alpha);
        SUF(nlcg_set_forces)(s, c, 0, 1);
        eta0 = eta;
        if (c->alpha * c->alpha * vdv < c->seceps * c->seceps) break;
        if ((c->sec > 1) && (c->afrac > 0) && (fabs(c->alpha) < fabs(c->dxsum)) &&
            (fabs(c->alpha) / fabs(c->dxsum) < c->afrac))
            break;
        if (c->stepmax > 0 && c->dxsum * c->dxsum * vdv > c->stepmax * c->stepmax) {
            c->k = 0;
            break;
        }
    }
    c->alphavmax = sqrt(c->alpha * c->alpha * vdv);
    double Kold = c->Knew;
    double Kmid = SUF(nlcg_dot)(s, c, 1);
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) {
            for (int k = 0; k < NDIM; k++) A[i].a[k] = 0;
            continue;
        }
        for (int k = 0; k < NDIM; k++) A[i].a[k] = A[i].f[k];
    }
    c->al = c->fl;
    c->Knew = SUF(nlcg_dot)(s, c, 1);
    c->beta = (c->Knew - Kmid) / Kold;
    c->betaused = c->beta;
    c->k++;
    if (c->k >= c->kmax || isinf(c->betaused) || isnan(c->betaused) || c->betaused <= 0) {
        c->k = 0;
        c->betaused = 0;
    } else if (c->betaused > 1) {
        c->betaused = 1;
    }
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) {
            for (int k = 0; k < NDIM; k++) A[i].v[k] = 0;
            continue;
        }
        for (int k = 0; k < NDIM; k++) A[i].v[k] = A[i].a[k] + A[i].v[k] * c->betaused;
    }
    c->vl = c->al + c->betaused * c->vl;
}

/* descend (src/collection.cpp:837-853) */
static void SUF(nlcg_descend)(or_system *s, or_nlcg *c) {
    SUF(nlcg_set_forces)(s, c, 0, 1);
    SUF(atom) *A = (SUF(atom) *)s->atoms;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) continue;
        for (int k = 0; k < NDIM; k++) A[i].v[k] = A[i].a[k] = A[i].f[k];
    }
    c->al = c->fl;
    c->vl = c->fl;
    SUF(nlcg_stepx)(s, c, c->dt);
    SUF(sys_update_list)(s, 0);
}

/* kinetic_energy (src/collection.cpp:575-590): masses ignored */
static double SUF(nlcg_kinetic)(const or_system *s, const or_nlcg *c) {
    const SUF(atom) *A = (const SUF(atom) *)s->atoms;
    double Lfac = exp(c->vl / (c->kappa * NDIM));
    double E = 0;
    for (int64_t i = 0; i < s->natoms; i++) {
        if (SUF(is_fixed)(A[i].m)) continue;
        double w[NDIM];
        for (int k = 0; k < NDIM; k++) w[k] = A[i].v[k] + (A[i].x[k] * Lfac);
        E += SUF(dot)(w, w);
    }
    return E / 2.0;
}
