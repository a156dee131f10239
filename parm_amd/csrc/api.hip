// api.hip -- C ABI of libparm_hip.so (declared in include/parm_hip.h).
//
// Host orchestration of the device path: owns the per-system device state
// (structure-of-arrays, spatially sorted slots), decides list rebuilds with
// the reference's skin rule, and runs the step loop.  No CPU compute path
// exists: every observable is computed on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/parm_hip.h"
#include "common.hpp"
#include "kernels.hpp"
#include "nlcg.hpp"
#include "pairtable.hpp"
#include "tiles.hpp"

using namespace parm;

namespace {
thread_local std::string g_err;
}  // namespace

namespace parm {
int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
}  // namespace parm

namespace {

int fail(int code, const std::string &msg) { return set_error(code, msg); }

#define PARM_HIP(call)                                                         \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return fail(PARM_E_DEVICE, std::string("HIP error ") +             \
                                           hipGetErrorString(e_) + " in " +    \
                                           #call);                             \
    } while (0)

#define PARM_CHECK(expr)                                                       \
    do {                                                                       \
        int r_ = (expr);                                                       \
        if (r_ != PARM_OK) return r_;                                          \
    } while (0)

template <typename T>
int dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(PARM_E_NOMEM, "device allocation of " +
                                      std::to_string(sizeof(T) * count) +
                                      " bytes failed");
    }
    return PARM_OK;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

constexpr int kFlagRing = 4;
// flag words: [0, kFlagRing) step ring, then the list-build overflow flag, a
// scratch word for reductions that raise no flag, and the on-demand skin test
// dynamic LDS of one tiled build workgroup, at most (the stencil's atoms)
constexpr size_t kTiledLdsMax = 56 * 1024;
constexpr int kSlotBuild = kFlagRing + 1, kSlotScratch = kFlagRing + 2,
              kSlotSkin = kFlagRing + 3;

struct parm_system {
    int device = 0;
    hipStream_t stream = nullptr;
    int ndim = 3, prec = kFP64;
    int n = 0;
    size_t pos_bytes = 32;  // sizeof(PosStore<D>::T)

    Box box{};
    bool box_set = false;

    // per-slot device arrays (slot order = spatial sort order)
    void *pos = nullptr, *pos_alt = nullptr, *xref = nullptr, *xwrap = nullptr;
    double *v[3] = {nullptr, nullptr, nullptr};
    double *a[3] = {nullptr, nullptr, nullptr};
    double *f[3] = {nullptr, nullptr, nullptr};
    double *m = nullptr, *diam = nullptr;
    int32_t *type = nullptr, *orig = nullptr, *rank = nullptr;
    // permutation targets: one spare per permuted column (PermuteSet order)
    void *pos_spare = nullptr;
    double *spare_d[kPermD] = {};
    int32_t *spare_i[kPermI] = {};
    // staging for uploads/downloads (n x ndim doubles, original order)
    double *stage = nullptr;
    double *stage_small = nullptr;  // per-atom scalars / int32 in original order

    // neighbour list
    bool list_cfg = false;
    double skin = 0.0, dmax = 0.0, dmin = 0.0;  // list diameters of the members
    int64_t nmem = 0;
    bool ignorechanged = true;
    uint32_t which = 0;
    int64_t npairs = 0;
    bool built = false;
    CellGrid grid{};
    int cap = 0, npad = 0, type_bits = 0;
    int tpa = 4;            // lanes per atom in the force kernel (autotuned)
    bool tpa_fixed = false, tuned = false;
    uint32_t *nbr = nullptr;       // untiled: uint32 entries; tiled: uint16 staged slots
    int32_t *ncount = nullptr;
    uint16_t *lcount = nullptr;   // per-lane entry counts, npad x 8
    // tiled layout (tiles.hpp): the current list is tiled when `tiled`
    bool tiled = false, tile_ok = false;
    int layout_pref = 0;  // 0: tiled when possible, until tuned; 1: tiled; 2: untiled
    int4 *tiles = nullptr;
    int2 *truns = nullptr;
    int tiles_cap = 0, ntiles = 0, smax = 0;
    size_t lds_max = 0;  // dynamic LDS a tiled workgroup may take
    int *tinfo = nullptr, *htinfo = nullptr;
    int spec_smax = 0;  // last plan's largest stencil: the next rebuild's speculative bound
    BuildReport *hrep = nullptr, *hrep_dev = nullptr;  // pinned build report
    uint32_t *keys = nullptr, *tkeys = nullptr, *hist = nullptr;
    int32_t *vals = nullptr, *tvals = nullptr, *cstart = nullptr;
    int cstart_cap = 0;
    unsigned long long *stats = nullptr, *bpart = nullptr;

    // NeighborList::ignore: excluded pairs as sorted keys (hi << 32 | lo) of
    // atom indices, and their device CSR (both directions, over original
    // atom indices) uploaded before the next build when changed
    std::vector<uint64_t> excl;
    bool excl_dirty = false;
    int32_t *ex_off = nullptr, *ex_part = nullptr;
    int64_t ex_part_cap = 0, ex_npart = 0;

    // pair interaction
    int pair_kind = kPairNone;
    int ntypes = 0;
    int single = 1, harmonic = 0;
    PairParam *tab = nullptr;
    PairParam tab0{};  // host copy of tab[0]

    // reductions and flags
    double *part = nullptr, *out = nullptr;
    uint32_t *fcnt = nullptr;   // the multi-workgroup skin reduction's counter
    double *fgpart = nullptr;   // and its per-workgroup results
    uint32_t *dflag = nullptr;      // flag words (see kSlot*)
    uint32_t *hflag = nullptr;      // pinned mirror, written by the kernels
    uint32_t *hflag_dev = nullptr;  // its device-side address
    double *hout = nullptr;         // pinned mirror of out
    hipEvent_t ev_flag = nullptr;
    hipEvent_t ev_ring[kFlagRing] = {};

    // asynchronous snapshots for output: a ring of device staging buffers
    // and pinned host copies, filled on the system's stream and copied to the
    // host on a stream of their own
    struct Snap {
        double *dev = nullptr, *host = nullptr;
        hipEvent_t ready = nullptr, done = nullptr;
        int64_t ticket = -1;
        int what = 0;
        double L[3] = {0, 0, 0};
    };
    static constexpr int kSnapRing = 4;
    Snap snap[kSnapRing];
    hipStream_t out_stream = nullptr;
    int64_t snap_next = 0;

    // instrumentation: event pairs around step-kernel launches (every
    // prof_every-th one: an event pair costs the stream a few microseconds)
    // and rebuilds
    bool prof = false;
    int prof_every = 1;
    int64_t prof_seq = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_step, prof_rebuild;

    AtomArrays arrays(bool alt_out = false) const {
        AtomArrays A{};
        A.n = n;
        A.pos_in = pos;
        A.pos_out = alt_out ? pos_alt : pos;
        for (int k = 0; k < 3; k++) {
            A.v[k] = v[k];
            A.a[k] = a[k];
            A.f[k] = f[k];
        }
        A.m = m;
        A.type = type;
        A.rank = rank;
        A.xref = xref;
        return A;
    }
    ListView listview() const {
        ListView L{};
        L.nbr = nbr;
        L.ncount = ncount;
        L.lcount = lcount;
        L.npad = npad;
        L.type_bits = type_bits;
        L.tpa = tpa;
        L.built = built ? 1 : 0;
        L.tiled = tiled ? 1 : 0;
        L.ntiles = ntiles;
        L.smax = smax;
        L.wcap = tpa > 0 ? cap / tpa : 0;
        L.tiles = tiles;
        L.truns = truns;
        L.keys = keys;
        L.cstart = cstart;
        L.grid = grid;
        return L;
    }
    PairView pairview() const {
        PairView P{};
        P.kind = pair_kind;
        P.ntypes = ntypes;
        P.single = single;
        P.harmonic = harmonic;
        P.tab = tab;
        P.p0 = tab0;
        return P;
    }
    ExclView exview() const {
        ExclView E{};
        E.orig = orig;
        E.off = ex_off;
        E.part = ex_part;
        E.any = ex_npart > 0 ? 1 : 0;
        return E;
    }
    ReduceBuf redbuf(int flag_index) const {
        ReduceBuf R{};
        R.part = part;
        R.pstride = force_blocks(std::max(n, 1), 8);
        R.out = out;
        R.flag = dflag + flag_index;
        R.skin = skin;
        R.hflag = hflag_dev + flag_index;
        R.cnt = fcnt;
        R.gpart = fgpart;
        return R;
    }
};

namespace {

int check_sys(const parm_system *s) {
    if (!s) return fail(PARM_E_INVALID, "null parm_system");
    return PARM_OK;
}

int set_device(const parm_system *s) {
    PARM_HIP(hipSetDevice(s->device));
    return PARM_OK;
}

void free_list(parm_system *s) {
    dfree(s->nbr);
    dfree(s->ncount);
    dfree(s->lcount);
    s->cap = 0;
}

// Cell grid for the current box / diameters / skin: edge >= largest list
// radius (dmax + skin) with a 1e-9 relative margin for rounding in the
// binning; fewer than 3 cells along a dimension collapses it to 1 cell.
void make_grid(parm_system *s) {
    CellGrid g{};
    const double rmax = (s->dmax + s->skin) * (1.0 + 1e-9);
    long total = 1;
    for (int k = 0; k < 3; k++) {
        g.nc[k] = 1;
        g.invL[k] = 0.0;
    }
    for (int k = 0; k < s->ndim; k++) {
        double q = rmax > 0 ? std::floor(s->box.L[k] / rmax) : 1.0;
        int nc = (q >= 3.0) ? (int)std::min(q, 4096.0) : 1;
        g.nc[k] = nc;
        g.invL[k] = s->box.invL[k];
    }
    // keep the grid within ~4 cells per atom (and at least 27 cells)
    for (;;) {
        total = 1;
        for (int k = 0; k < s->ndim; k++) total *= g.nc[k];
        if (total <= std::max<long>(4L * s->n, 27L)) break;
        int kk = 0;
        for (int k = 1; k < s->ndim; k++)
            if (g.nc[k] > g.nc[kk]) kk = k;
        g.nc[kk] = g.nc[kk] / 2 >= 3 ? g.nc[kk] / 2 : 1;
    }
    g.ncell = (int)total;
    s->grid = g;
}

// Per-lane entry counts are 16-bit (lcount): a list whose capacity could
// give one lane more than 65535 entries is refused, not wrapped.
int check_lane_counts(int64_t cap, int tpa) {
    if (cap > (int64_t)65535 * tpa)
        return fail(PARM_E_UNSUPPORTED,
                    "more than 65535 list entries per lane (capacity " + std::to_string(cap) +
                        " at " + std::to_string(tpa) +
                        " lanes per atom): use more lanes per atom or a smaller skin");
    return PARM_OK;
}
int check_lane_counts(const parm_system *s) { return check_lane_counts(s->cap, s->tpa); }

int ensure_list_capacity(parm_system *s, int cap) {
    if (cap <= s->cap && s->nbr) return PARM_OK;
    // refused before the list is reallocated
    PARM_CHECK(check_lane_counts(((int64_t)cap + 7) & ~(int64_t)7, s->tpa));
    free_list(s);
    s->cap = (cap + 7) & ~7;
    PARM_CHECK(dalloc(&s->nbr, (size_t)s->cap * s->npad));
    PARM_CHECK(dalloc(&s->ncount, (size_t)s->npad));
    PARM_CHECK(dalloc(&s->lcount, (size_t)s->npad * 8));
    return PARM_OK;
}

int initial_capacity(const parm_system *s) {
    double vol = 1.0;
    for (int k = 0; k < s->ndim; k++) vol *= s->box.L[k];
    const double r = s->dmax + s->skin;
    const double ball = s->ndim == 3 ? 4.0 / 3.0 * M_PI * r * r * r : M_PI * r * r;
    double mean = vol > 0 ? (double)s->nmem / vol * ball : 0.0;
    double cap = 1.35 * mean + 24.0;
    if (cap > (double)s->nmem) cap = (double)std::max<int64_t>(s->nmem, 8);
    return std::max(16, (int)cap);
}

// Permute every per-slot array by perm (new slot i <- old slot perm[i]).
// Every per-slot array into the new slot order, in one launch: the position
// buffers (the alternate one holds the next step's drifted positions inside
// a fused step loop; xref is free to take one of them, the list build
// rewrites it), the ndim components of v, a, f, the masses and diameters,
// and the type / original index / member rank.  Each lands in its spare,
// which then becomes the column.  keep_af = false leaves a and f in the old
// order: for a rebuild inside the step loop whose next launch writes both
// before anything reads them.
int permute_all(parm_system *s, const int32_t *perm, bool keep_af = true) {
    PermuteSet P{};
    P.pos_src[0] = s->pos;
    P.pos_dst[0] = s->xref;
    P.pos_src[1] = s->pos_alt;
    P.pos_dst[1] = s->pos_spare;
    double **dcols[kPermD];
    int nd = 0;
    for (int k = 0; k < s->ndim; k++) dcols[nd++] = &s->v[k];
    if (keep_af) {
        for (double **vec : {s->a, s->f})
            for (int k = 0; k < s->ndim; k++) dcols[nd++] = &vec[k];
    }
    dcols[nd++] = &s->m;
    dcols[nd++] = &s->diam;
    for (int c = 0; c < nd; c++) {
        P.d_src[c] = *dcols[c];
        P.d_dst[c] = s->spare_d[c];
    }
    P.nd = nd;
    int32_t **icols[kPermI] = {&s->type, &s->orig, &s->rank};
    for (int c = 0; c < kPermI; c++) {
        P.i_src[c] = *icols[c];
        P.i_dst[c] = s->spare_i[c];
    }
    P.ni = kPermI;
    PARM_HIP(launch_permute(s->ndim, s->n, perm, P, s->stream));
    std::swap(s->pos, s->xref);
    std::swap(s->pos_alt, s->pos_spare);
    for (int c = 0; c < nd; c++) std::swap(*dcols[c], s->spare_d[c]);
    for (int c = 0; c < kPermI; c++) std::swap(*icols[c], s->spare_i[c]);
    return PARM_OK;
}

// Tiles of the current cell order (tiles.hpp): launch_tile_plan queues the
// plan (tile records, stencil runs, info = {tiles, largest stencil, bad});
// accept_tiles decides the layout from it.  The tiled layout is used when
// every tile's stencil fits in LDS and the tile count fits the reduction
// scratch; otherwise (or with PARM_TILED=0) the untiled layout.
int launch_tile_plan(parm_system *s) {
    if (s->n == 0) return PARM_OK;
    const int nrows = s->grid.ncell / s->grid.nc[0];
    const int need = 2 * (s->n / 64) + nrows + 4;
    if (need > s->tiles_cap) {
        dfree(s->tiles);
        dfree(s->truns);
        s->tiles_cap = need;
        PARM_CHECK(dalloc(&s->tiles, (size_t)need));
        PARM_CHECK(dalloc(&s->truns, (size_t)need * kTileRuns));
    }
    PARM_HIP(launch_tiles(s->ndim, s->n, s->grid, s->keys, s->cstart, s->tiles_cap, s->tiles,
                          s->truns, s->tinfo, s->stream));
    return PARM_OK;
}

// The tiled force kernel's staged stencil fits the LDS it may take (the
// list build's staging is smaller: a plan the build accepts can still need
// the untiled force layout).
bool tiled_force_fits(const parm_system *s) {
    return tiled_force_lds(s->ndim, s->prec, s->pair_kind, s->single != 0, s->smax) <=
           s->lds_max;
}

void accept_tiles(parm_system *s, int nt, int smax, int bad) {
    s->tiled = false;
    s->tile_ok = false;
    s->ntiles = 0;
    s->smax = 0;
    const int pstride = force_blocks(std::max(s->n, 1), 8);
    if (s->n == 0 || bad || nt <= 0 || nt > pstride || tiled_build_lds(smax) > s->lds_max)
        return;
    s->tile_ok = true;
    s->ntiles = nt;
    s->smax = std::max(smax, 1);
    s->tiled = s->layout_pref != 2 && tiled_force_fits(s);
}

// The plan queued by launch_tile_plan, read back (one host round trip).
int plan_tiles(parm_system *s) {
    if (s->n == 0) {
        accept_tiles(s, 0, 0, 0);
        return PARM_OK;
    }
    PARM_HIP(hipMemcpyAsync(s->htinfo, s->tinfo, 3 * sizeof(int), hipMemcpyDeviceToHost,
                            s->stream));
    PARM_HIP(hipStreamSynchronize(s->stream));
    accept_tiles(s, s->htinfo[0], s->htinfo[1], s->htinfo[2]);
    return PARM_OK;
}

// Builds the list at positions `at` (the current ones, or xref for a
// relayout) in the layout chosen by accept_tiles.
int build_list(parm_system *s, const void *at) {
    PARM_CHECK(check_lane_counts(s));
    hipStream_t st = s->stream;
    if (s->tile_ok) {
        // the tiled build writes either list layout
        PARM_HIP(launch_build_tiled(s->ndim, s->n, s->ntiles, s->smax, s->tiles, s->truns, at,
                                    s->xwrap,
                                    s->rank, s->diam, s->type, s->type_bits, s->keys,
                                    s->cstart, s->grid, s->box, s->skin, s->dmax, s->tpa,
                                    s->cap, s->npad, !s->tiled, s->nbr, s->ncount, s->lcount,
                                    s->xref, s->bpart, s->stats, s->dflag + kSlotBuild,
                                    nullptr, s->dmin == s->dmax, s->exview(), s->hrep_dev,
                                    st));
    } else {
        PARM_HIP(launch_build_list(s->ndim, s->n, at, s->xwrap, s->rank, s->diam, s->type,
                                   s->keys, s->cstart, s->grid, s->box, s->skin, s->dmax,
                                   s->type_bits, s->tpa, s->cap, s->npad, s->nbr, s->ncount,
                                   s->lcount, s->xref, s->bpart, s->stats, s->dflag + kSlotBuild,
                                   s->exview(), s->hrep_dev, st));
    }
    return PARM_OK;
}

// The report of the build just queued, once the stream has finished.
int read_report(parm_system *s, BuildReport *r) {
    PARM_HIP(hipStreamSynchronize(s->stream));
    const volatile BuildReport *h = s->hrep;
    r->sum = h->sum;
    r->max = h->max;
    r->ntiles = h->ntiles;
    r->smax = h->smax;
    r->bad = h->bad;
    r->flag = h->flag;
    return PARM_OK;
}

// build_list until the list fits: the capacity a layout needs is the
// fullest lane's count times the lanes per atom, so it can grow on a rebuild
// and on a change of lanes per atom alike.  Sets npairs.
int build_checked(parm_system *s, const void *at) {
    for (int attempt = 0; attempt < 4; attempt++) {
        PARM_CHECK(build_list(s, at));
        BuildReport r{};
        if (s->n == 0) {
            s->npairs = 0;
            return PARM_OK;
        }
        PARM_CHECK(read_report(s, &r));
        if (r.max <= (unsigned long long)s->cap) {
            s->npairs = (int64_t)(r.sum / 2);
            return PARM_OK;
        }
        PARM_CHECK(ensure_list_capacity(s, (int)(r.max + r.max / 8 + 8)));
    }
    return fail(PARM_E_DEVICE, "neighbour list build did not converge");
}

// Rebuild-time list build with one host round trip: the tiled build is
// queued behind the tile plan before the host has read it, with the
// previous plan's largest stencil (plus a margin) as its LDS bound and the
// tile capacity as its grid.  The report then carries the plan: when the
// plan fits the bounds (and the list its capacity) the build stands;
// otherwise the plan is accepted as read and the list built again.
int build_speculative(parm_system *s) {
    PARM_CHECK(check_lane_counts(s));
    const int pstride = force_blocks(std::max(s->n, 1), 8);
    const int grid = std::min(s->tiles_cap, pstride);
    // the largest stencil whose staging fits the LDS a workgroup may take
    const size_t fixed = tiled_build_lds(0), per = tiled_build_lds(1) - fixed;
    const int limit = s->lds_max > fixed ? (int)((s->lds_max - fixed) / per) : 0;
    const int bound = std::min(limit, s->spec_smax + std::max(s->spec_smax / 8, 64));
    // the layout the force kernel last took (accept_tiles re-decides it
    // from the plan as read; a mismatch rebuilds)
    const bool tiled = s->tiled;
    PARM_HIP(launch_build_tiled(s->ndim, s->n, grid, bound, s->tiles, s->truns, s->pos, s->xwrap,
                                s->rank,
                                s->diam, s->type, s->type_bits, s->keys, s->cstart, s->grid,
                                s->box, s->skin, s->dmax, s->tpa, s->cap, s->npad, !tiled,
                                s->nbr, s->ncount, s->lcount, s->xref, s->bpart, s->stats,
                                s->dflag + kSlotBuild, s->tinfo, s->dmin == s->dmax,
                                s->exview(), s->hrep_dev, s->stream));
    BuildReport r{};
    PARM_CHECK(read_report(s, &r));
    accept_tiles(s, r.ntiles, r.smax, r.bad);
    s->spec_smax = s->tile_ok ? s->smax : 0;
    const bool fits = s->tile_ok && s->tiled == tiled && r.ntiles <= grid &&
                      !(r.flag & kFlagReplan);
    if (fits && r.max <= (unsigned long long)s->cap) {
        s->npairs = (int64_t)(r.sum / 2);
        return PARM_OK;
    }
    if (fits) PARM_CHECK(ensure_list_capacity(s, (int)(r.max + r.max / 8 + 8)));
    return build_checked(s, s->pos);
}

// The exclusion CSR of s->excl on the device: for every atom (original
// index) its excluded partners, sorted; a pair of an atom with itself never
// reaches the predicate and is left out.
int upload_exclusions(parm_system *s) {
    if (!s->excl_dirty) return PARM_OK;
    const int n = s->n;
    std::vector<int32_t> off((size_t)n + 1, 0);
    for (uint64_t k : s->excl) {
        const int32_t hi = (int32_t)(k >> 32), lo = (int32_t)(k & 0xffffffffu);
        if (hi == lo) continue;
        off[(size_t)hi + 1]++;
        off[(size_t)lo + 1]++;
    }
    for (int i = 0; i < n; i++) off[(size_t)i + 1] += off[i];
    const int64_t np = off[n];
    std::vector<int32_t> part((size_t)std::max<int64_t>(np, 1));
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (uint64_t k : s->excl) {
        const int32_t hi = (int32_t)(k >> 32), lo = (int32_t)(k & 0xffffffffu);
        if (hi == lo) continue;
        part[fill[hi]++] = lo;
        part[fill[lo]++] = hi;
    }
    for (int i = 0; i < n; i++) std::sort(part.begin() + off[i], part.begin() + off[i + 1]);
    if (!s->ex_off) PARM_CHECK(dalloc(&s->ex_off, (size_t)n + 1));
    if (np > s->ex_part_cap) {
        dfree(s->ex_part);
        s->ex_part_cap = np + np / 2 + 16;
        PARM_CHECK(dalloc(&s->ex_part, (size_t)s->ex_part_cap));
    }
    if (np > 0) {
        PARM_HIP(hipMemcpyAsync(s->ex_off, off.data(), sizeof(int32_t) * off.size(),
                                hipMemcpyHostToDevice, s->stream));
        PARM_HIP(hipMemcpyAsync(s->ex_part, part.data(), sizeof(int32_t) * np,
                                hipMemcpyHostToDevice, s->stream));
        PARM_HIP(hipStreamSynchronize(s->stream));
    }
    s->ex_npart = np;
    s->excl_dirty = false;
    return PARM_OK;
}

// NeighborList::update_list's rebuild branch (src/trackers.cpp:55-69):
// lastlocs = x, then the full pair list from the predicate.
int rebuild(parm_system *s, bool keep_af = true) {
    if (!s->list_cfg) return fail(PARM_E_STATE, "neighbour list not configured");
    PARM_CHECK(upload_exclusions(s));
    hipStream_t st = s->stream;
    const int n = s->n;
    make_grid(s);
    if (s->grid.ncell + 1 > s->cstart_cap) {
        dfree(s->cstart);
        s->cstart_cap = s->grid.ncell + 1;
        PARM_CHECK(dalloc(&s->cstart, (size_t)s->cstart_cap));
    }
    // from here on the previous list no longer matches the atom order
    s->built = false;
    if (n > 0) {
        PARM_HIP(launch_cell_keys(s->ndim, n, s->pos, s->rank, s->grid, s->keys, s->vals,
                                  st));
        int bits = 1;
        while ((1u << bits) <= (uint32_t)s->grid.ncell && bits < 32) bits++;
        PARM_HIP(radix_sort(n, bits, s->keys, s->vals, s->tkeys, s->tvals, s->hist, st));
        PARM_CHECK(permute_all(s, s->vals, keep_af));
    }
    PARM_HIP(launch_cell_bounds(n, s->grid.ncell, s->keys, s->cstart, st));
    if (s->cap == 0) PARM_CHECK(ensure_list_capacity(s, initial_capacity(s)));
    PARM_CHECK(launch_tile_plan(s));
    if (n > 0 && s->spec_smax > 0) {
        PARM_CHECK(build_speculative(s));
    } else {
        PARM_CHECK(plan_tiles(s));
        PARM_CHECK(build_checked(s, s->pos));
        s->spec_smax = s->tile_ok ? s->smax : 0;
    }
    s->built = true;
    s->which++;
    s->ignorechanged = false;
    return PARM_OK;
}

// Flag slot idx once the stream has finished: the skin-test launches store
// it in the pinned mirror themselves (nothing ran for an empty system).
int read_flag(parm_system *s, int idx, uint32_t *flag) {
    PARM_HIP(hipStreamSynchronize(s->stream));
    *flag = s->n > 0 ? *(volatile uint32_t *)(s->hflag + idx) : 0u;
    return PARM_OK;
}

int update_list(parm_system *s, bool force, int *rebuilt) {
    if (rebuilt) *rebuilt = 0;
    if (!s->list_cfg) return fail(PARM_E_STATE, "neighbour list not configured");
    if (!force && !s->ignorechanged) {
        PARM_HIP(launch_skin(s->ndim, s->arrays(), s->redbuf(kSlotSkin), s->stream));
        uint32_t fl = 0;
        PARM_CHECK(read_flag(s, kSlotSkin, &fl));
        if (fl & kFlagNonFinite)
            return fail(PARM_E_NONFINITE, "non-finite position in skin test");
        if (!(fl & kFlagRebuild)) return PARM_OK;
    }
    PARM_CHECK(rebuild(s));
    if (rebuilt) *rebuilt = 1;
    return PARM_OK;
}

int require_ready(parm_system *s) {
    if (!s->box_set) return fail(PARM_E_STATE, "box not set (parm_set_box)");
    if (!s->list_cfg) return fail(PARM_E_STATE, "neighbour list not set (parm_nlist_set)");
    if (s->pair_kind == kPairNone)
        return fail(PARM_E_STATE, "no pair interaction set");
    return PARM_OK;
}

int set_forces(parm_system *s, bool and_a) {
    StepCoef c{};
    PARM_HIP(launch_force(s->ndim, s->prec, and_a ? kStageForceA : kStageForceOnly, false,
                          s->arrays(), s->listview(), s->pairview(), s->box, c,
                          s->redbuf(kSlotScratch), s->stream));
    return PARM_OK;
}

hipEvent_t pool_event(parm_system *s) {
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        s->ev_pool.push_back(e);
    }
    return s->ev_pool[s->ev_used++];
}

// Brackets a launch with timing events when instrumentation is on.  A force
// launch given end() as its ReduceBuf::done records the end itself, right
// after the force kernel (its k_finalize then falls outside the bracket).
struct ProfScope {
    parm_system *s;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> *dst;
    hipEvent_t a = nullptr, b = nullptr;
    bool callee_end = false;
    ProfScope(parm_system *s_, std::vector<std::pair<hipEvent_t, hipEvent_t>> *d,
              bool sample = true)
        : s(s_), dst(d) {
        if (!s->prof || !sample) return;
        a = pool_event(s);
        b = pool_event(s);
        if (a && b) (void)hipEventRecord(a, s->stream);
    }
    hipEvent_t end() {
        if (!s->prof || !a || !b) return nullptr;
        callee_end = true;
        return b;
    }
    ~ProfScope() {
        if (!s->prof || !a || !b) return;
        if (!callee_end) (void)hipEventRecord(b, s->stream);
        dst->emplace_back(a, b);
    }
};

// Rewrites the current list in the layout of s->tpa without changing its
// pairs: the build runs again on the positions of the last rebuild (xref),
// with the cell keys and bounds of that rebuild.
int relayout(parm_system *s) {
    if (!s->built || s->n == 0) return PARM_OK;
    if (s->tiled && !(s->tile_ok && tiled_force_fits(s))) s->tiled = false;
    PARM_CHECK(build_checked(s, s->xref));
    return PARM_OK;
}

// Autotuner of the step kernel: list layout (tiled / untiled, tiles.hpp) x
// lanes per atom.  Each candidate lays the list out for itself (same pairs:
// relayout) and times three launches of the run's own fused step kernel
// (`stage`), with its velocity / acceleration / force outputs redirected to
// scratch copies and its position output in pos_alt (which the first real
// step overwrites), so the state is untouched.  The fastest is kept for the
// life of the system; PARM_TPA / PARM_TILED pin either choice.
int autotune(parm_system *s, int stage, const StepCoef &c) {
    if (s->tuned || !s->built || s->n == 0) return PARM_OK;
    std::vector<std::pair<int, int>> cands;  // (tiled, tpa)
    std::vector<int> layouts;
    const bool tiled_ok = s->tile_ok && tiled_force_fits(s);
    if (tiled_ok && s->layout_pref != 2) layouts.push_back(1);
    if (s->layout_pref != 1 || !tiled_ok) layouts.push_back(0);
    for (int lay : layouts) {
        if (s->tpa_fixed) {
            cands.emplace_back(lay, s->tpa);
        } else {
            for (int tpa : {1, 2, 4, 8}) cands.emplace_back(lay, tpa);
        }
    }
    if (cands.size() <= 1) {
        if (!cands.empty() && (cands[0].first != (int)s->tiled || cands[0].second != s->tpa)) {
            s->tiled = cands[0].first != 0;
            s->tpa = cands[0].second;
            PARM_CHECK(relayout(s));
        }
        s->tuned = true;
        return PARM_OK;
    }
    hipStream_t st = s->stream;
    const size_t n = (size_t)s->n;
    double *sc = nullptr;
    PARM_CHECK(dalloc(&sc, 9 * n));
    AtomArrays A = s->arrays(true);
    for (int k = 0; k < 3; k++) {
        A.v[k] = sc + k * n;
        A.a[k] = sc + (3 + k) * n;
        A.f[k] = sc + (6 + k) * n;
    }
    int rc = PARM_OK;
    for (int k = 0; k < s->ndim && rc == PARM_OK; k++) {
        if (hipMemcpyAsync(A.v[k], s->v[k], n * sizeof(double), hipMemcpyDeviceToDevice, st) !=
                hipSuccess ||
            hipMemcpyAsync(A.a[k], s->a[k], n * sizeof(double), hipMemcpyDeviceToDevice, st) !=
                hipSuccess)
            rc = fail(PARM_E_DEVICE, "autotune copy failed");
    }
    hipEvent_t e0 = pool_event(s), e1 = pool_event(s);
    if (rc == PARM_OK && (!e0 || !e1)) rc = fail(PARM_E_DEVICE, "event creation failed");
    size_t best = 0;
    float best_ms = 1e30f;
    for (size_t q = 0; q < cands.size() && rc == PARM_OK; q++) {
        s->tiled = cands[q].first != 0;
        s->tpa = cands[q].second;
        if ((rc = relayout(s)) != PARM_OK) break;
        float tmin = 1e30f;
        for (int r = 0; r < 4 && rc == PARM_OK; r++) {  // warm-up + best of 3
            hipError_t e = hipEventRecord(e0, st);
            if (e == hipSuccess)
                e = launch_force(s->ndim, s->prec, stage, false, A, s->listview(),
                                 s->pairview(), s->box, c, s->redbuf(kSlotScratch), st);
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                rc = fail(PARM_E_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
                break;
            }
            if (r > 0) tmin = std::min(tmin, ms);
        }
        if (tmin < best_ms * 0.97f) {  // prefer the earlier candidate on ties
            best_ms = tmin;
            best = q;
        }
    }
    if (e0 && e1) s->ev_used -= 2;
    (void)hipStreamSynchronize(st);
    dfree(sc);
    if (rc != PARM_OK) return rc;
    s->tiled = cands[best].first != 0;
    s->tpa = cands[best].second;
    s->layout_pref = s->tiled ? 1 : 2;
    PARM_CHECK(relayout(s));
    s->tuned = true;
    return PARM_OK;
}

// Records the ring event of flag slot q: the launch just queued stores the
// slot in the pinned mirror (an empty system launches nothing: the slot is
// cleared here).
int post_flag(parm_system *s, int slot) {
    if (s->n == 0) s->hflag[slot] = 0;
    PARM_HIP(hipEventRecord(s->ev_ring[slot], s->stream));
    return PARM_OK;
}

// nsteps x timestep() of CollectionVerlet / CollectionDamped, fused.
//
// The reference step is: drift (first loop), set_forces, kick (second loop),
// update_trackers (skin test -> maybe rebuild).  Here one kernel per step does
// the force + kick of step q AND the drift of step q+1 into the alternate
// position buffer, plus the skin test of those drifted positions.  The list
// decision for the end of step q needs the skin test of x_{q+1}, which the
// previous kernel produced: the host waits for that flag while the current
// kernel runs.  A rebuild at x_{q+1} permutes both position buffers and
// re-tests x_{q+2} against the new reference positions.  The last step runs
// force + kick only, so the state returned is at a full step.
int run_steps(parm_system *s, bool damped, const StepCoef &c, int64_t nsteps) {
    PARM_CHECK(require_ready(s));
    if (nsteps <= 0) return PARM_OK;
    const int kick = damped ? kStageDampedKick : kStageVerletKick;
    const int kick_drift = damped ? kStageDampedKickDrift : kStageVerletKickDrift;
    if (nsteps >= 16) PARM_CHECK(autotune(s, kick_drift, c));
    hipStream_t st = s->stream;
    // first half of step 0
    PARM_HIP(launch_drift(s->ndim, kick, s->arrays(), c, s->redbuf(0), st));
    PARM_CHECK(post_flag(s, 0));
    for (int64_t q = 0; q < nsteps; q++) {
        const bool last = (q == nsteps - 1);
        const int slot = (int)(q % kFlagRing), next = (int)((q + 1) % kFlagRing);
        if (!last) {
            {
                ProfScope ps(s, &s->prof_step, s->prof_seq++ % s->prof_every == 0);
                ReduceBuf R = s->redbuf(next);
                R.done = ps.end();
                PARM_HIP(launch_force(s->ndim, s->prec, kick_drift, false,
                                      s->arrays(true), s->listview(), s->pairview(),
                                      s->box, c, R, st));
            }
            PARM_CHECK(post_flag(s, next));
        } else {
            PARM_HIP(launch_force(s->ndim, s->prec, kick, false, s->arrays(),
                                  s->listview(), s->pairview(), s->box, c,
                                  s->redbuf(kSlotScratch), st));
        }
        // update_trackers at the end of step q: skin test of x_{q+1}
        PARM_HIP(hipEventSynchronize(s->ev_ring[slot]));
        const uint32_t fl = *(volatile uint32_t *)(s->hflag + slot);
        if (fl & kFlagNonFinite)
            return fail(PARM_E_NONFINITE,
                        "position became non-finite at step " + std::to_string(q));
        if ((fl & kFlagRebuild) || s->ignorechanged) {
            {
                ProfScope ps(s, &s->prof_rebuild);
                // before the last step's end the next launch rewrites a and
                // f before anything reads them: they need no permutation
                PARM_CHECK(rebuild(s, last));
            }
            if (!last) {
                AtomArrays A = s->arrays();
                A.pos_in = s->pos_alt;
                PARM_HIP(launch_skin(s->ndim, A, s->redbuf(next), st));
                PARM_CHECK(post_flag(s, next));
            }
        }
        if (!last) std::swap(s->pos, s->pos_alt);
    }
    return PARM_OK;
}

int upload_vec(parm_system *s, int what, const double *host) {
    const size_t bytes = sizeof(double) * (size_t)s->n * s->ndim;
    if (bytes == 0) return PARM_OK;
    PARM_HIP(hipMemcpyAsync(s->stage, host, bytes, hipMemcpyHostToDevice, s->stream));
    PARM_HIP(launch_upload(s->ndim, what, s->arrays(), s->orig, s->stage, s->stream));
    PARM_HIP(hipStreamSynchronize(s->stream));
    return PARM_OK;
}

int download_vec(parm_system *s, int what, double *host) {
    const size_t bytes = sizeof(double) * (size_t)s->n * s->ndim;
    if (bytes == 0) return PARM_OK;
    PARM_HIP(launch_download(s->ndim, what, s->arrays(), s->orig, s->stage, s->stream));
    PARM_HIP(hipMemcpyAsync(host, s->stage, bytes, hipMemcpyDeviceToHost, s->stream));
    PARM_HIP(hipStreamSynchronize(s->stream));
    return PARM_OK;
}

// Per-atom scalar in original order -> slot order.
template <typename T>
int upload_scalar(parm_system *s, const T *host, T *dst) {
    if (s->n == 0) return PARM_OK;
    static_assert(sizeof(T) <= sizeof(double), "scalar");
    PARM_HIP(hipMemcpyAsync(s->stage_small, host, sizeof(T) * s->n,
                            hipMemcpyHostToDevice, s->stream));
    PARM_HIP(gather<T>(s->n, s->orig, (const T *)s->stage_small, dst, s->stream));
    PARM_HIP(hipStreamSynchronize(s->stream));
    return PARM_OK;
}

// The reduction results of the launch just queued.  An empty system queues
// no reduction: its sums are exactly zero (the device buffer is not read).
int read_out(parm_system *s) {
    if (s->n == 0) {
        for (int k = 0; k < kRedWidth; k++) s->hout[k] = 0.0;
        return PARM_OK;
    }
    PARM_HIP(hipMemcpyAsync(s->hout, s->out, sizeof(double) * kRedWidth,
                            hipMemcpyDeviceToHost, s->stream));
    PARM_HIP(hipStreamSynchronize(s->stream));
    return PARM_OK;
}

void destroy(parm_system *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    dfree(s->pos);
    dfree(s->pos_alt);
    dfree(s->xref);
    dfree(s->xwrap);
    for (int k = 0; k < 3; k++) {
        dfree(s->v[k]);
        dfree(s->a[k]);
        dfree(s->f[k]);
    }
    dfree(s->m);
    dfree(s->diam);
    dfree(s->type);
    dfree(s->orig);
    dfree(s->rank);
    dfree(s->pos_spare);
    for (double *&p : s->spare_d) dfree(p);
    for (int32_t *&p : s->spare_i) dfree(p);
    dfree(s->stage);
    dfree(s->stage_small);
    free_list(s);
    dfree(s->keys);
    dfree(s->tkeys);
    dfree(s->hist);
    dfree(s->vals);
    dfree(s->tvals);
    dfree(s->cstart);
    dfree(s->tiles);
    dfree(s->truns);
    dfree(s->tinfo);
    if (s->htinfo) (void)hipHostFree(s->htinfo);
    dfree(s->stats);
    dfree(s->bpart);
    dfree(s->tab);
    dfree(s->ex_off);
    dfree(s->ex_part);
    dfree(s->part);
    dfree(s->out);
    dfree(s->fcnt);
    dfree(s->fgpart);
    dfree(s->dflag);
    if (s->hflag) (void)hipHostFree(s->hflag);
    if (s->hrep) (void)hipHostFree(s->hrep);
    if (s->hout) (void)hipHostFree(s->hout);
    for (parm_system::Snap &q : s->snap) {
        dfree(q.dev);
        if (q.host) (void)hipHostFree(q.host);
        if (q.ready) (void)hipEventDestroy(q.ready);
        if (q.done) (void)hipEventDestroy(q.done);
    }
    if (s->out_stream) {
        (void)hipStreamSynchronize(s->out_stream);
        (void)hipStreamDestroy(s->out_stream);
    }
    if (s->ev_flag) (void)hipEventDestroy(s->ev_flag);
    for (hipEvent_t &e : s->ev_ring)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

__global__ void k_iota(int n, int32_t *p) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = i;
}
__global__ void k_fill_i32(int n, int32_t *p, int32_t v) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int relayout(parm_system *s);

// OriginBox(Vec) / resize_to (src/box.cpp:17-20): edges L[ndim] > 0.
int apply_box(parm_system *s, const double *L) {
    for (int k = 0; k < s->ndim; k++)
        if (!(L[k] > 0) || !std::isfinite(L[k]))
            return fail(PARM_E_INVALID, "box edges must be finite and > 0");
    for (int k = 0; k < 3; k++) {
        double Lk = k < s->ndim ? L[k] : 1.0;
        s->box.L[k] = Lk;
        s->box.invL[k] = 1.0 / Lk;
        s->box.halfL[k] = Lk / 2;
    }
    s->box_set = true;
    return PARM_OK;
}

int build_pair_table(parm_system *s, int form, int ntypes, const double *eps,
                     const double *sigma, const double *third, const int32_t *atype) {
    if (ntypes < 1 || ntypes > 64)
        return fail(PARM_E_UNSUPPORTED, "ntypes must be in [1, 64]");
    if (!eps || !sigma || (!third && form != kFormLJRepulse) || (!atype && s->n > 0))
        return fail(PARM_E_INVALID, "null parameter array");
    for (int i = 0; i < s->n; i++)
        if (atype[i] < 0 || atype[i] >= ntypes)
            return fail(PARM_E_INVALID, "atom type out of range at atom " + std::to_string(i));
    const std::string asym = check_rows_symmetric(form, ntypes, eps, sigma);
    if (!asym.empty())
        return fail(PARM_E_UNSUPPORTED,
                    asym + ": the device evaluates each pair from both of its atoms");
    const int kind = form_kernel_kind(form);
    std::vector<PairParam> tab;
    bool harm = true;
    make_pair_table(form, ntypes, eps, sigma, third, tab, harm);
    PARM_CHECK(set_device(s));
    dfree(s->tab);
    PARM_CHECK(dalloc(&s->tab, tab.size()));
    PARM_HIP(hipMemcpy(s->tab, tab.data(), sizeof(PairParam) * tab.size(),
                       hipMemcpyHostToDevice));
    s->tab0 = tab[0];
    if (s->n > 0) PARM_CHECK(upload_scalar<int32_t>(s, atype, s->type));
    int bits = 0;
    while ((1 << bits) < ntypes) bits++;
    if (bits > 0 && (int64_t)s->n >= (int64_t(1) << (32 - bits)))
        return fail(PARM_E_UNSUPPORTED, "too many atoms for this many types");
    s->type_bits = bits;
    s->ntypes = ntypes;
    s->single = ntypes == 1 ? 1 : 0;
    s->harmonic = (kind == kPairRepulsion && harm) ? 1 : 0;
    s->pair_kind = kind;
    // List entries carry the neighbour's type: re-encode the current list
    // (same pairs, as NListed::update_pairs re-pairs the current list).
    PARM_CHECK(relayout(s));
    return PARM_OK;
}

}  // namespace

// ============================================================== the C ABI
extern "C" {

const char *parm_last_error(void) { return g_err.c_str(); }

int parm_version(int *major, int *minor) {
    if (major) *major = 1;
    if (minor) *minor = 0;
    return PARM_OK;
}

int parm_device_count(int *count) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    if (count) *count = c;
    return PARM_OK;
}

int parm_system_create(int ndim, int precision, int64_t natoms, const double *masses,
                       parm_system **out) {
    if (!out) return fail(PARM_E_INVALID, "out is null");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return fail(PARM_E_INVALID, "ndim must be 2 or 3");
    if (precision != PARM_FP64 && precision != PARM_MIXED)
        return fail(PARM_E_INVALID, "precision must be PARM_FP64 or PARM_MIXED");
    if (natoms < 0 || natoms >= (int64_t(1) << 31))
        return fail(PARM_E_INVALID, "natoms out of range");
    if (natoms > 0 && !masses) return fail(PARM_E_INVALID, "masses is null");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return fail(PARM_E_DEVICE, "no HIP device available (the device path has no CPU fallback)");
    }
    parm_system *s = new parm_system();
    s->ndim = ndim;
    s->prec = precision;
    s->n = (int)natoms;
    s->npad = (int)((natoms + 63) / 64 * 64);
    if (s->npad == 0) s->npad = 64;
    s->pos_bytes = ndim == 3 ? sizeof(double4) : sizeof(double2);
    if (const char *e = std::getenv("PARM_TPA")) {
        int t = std::atoi(e);
        if (t == 1 || t == 2 || t == 4 || t == 8) {
            s->tpa = t;
            s->tpa_fixed = true;
        }
    }
    if (const char *e = std::getenv("PARM_TILED")) s->layout_pref = std::atoi(e) == 0 ? 2 : 1;
    int rc = PARM_OK;
    auto bail = [&](int r) {
        destroy(s);
        return r;
    };
    if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(PARM_E_DEVICE, "hipGetDevice failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, s->device) != hipSuccess)
        return bail(fail(PARM_E_DEVICE, "hipGetDeviceProperties failed"));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return bail(fail(PARM_E_DEVICE, std::string("device is ") + prop.gcnArchName +
                                            "; this library is built for gfx950 only"));
    // leave room for the static LDS of the tiled kernels (stencil layout,
    // reduction scratch)
    s->lds_max = std::min(kTiledLdsMax, prop.sharedMemPerBlock > 8192
                                            ? (size_t)prop.sharedMemPerBlock - 8192
                                            : (size_t)0);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PARM_E_DEVICE, "stream creation failed"));
    if (hipEventCreateWithFlags(&s->ev_flag, hipEventDisableTiming) != hipSuccess)
        return bail(fail(PARM_E_DEVICE, "event creation failed"));
    for (hipEvent_t &e : s->ev_ring)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(PARM_E_DEVICE, "event creation failed"));
    const size_t n = (size_t)std::max(1, s->n);
    if ((rc = dalloc((char **)&s->pos, n * s->pos_bytes)) ||
        (rc = dalloc((char **)&s->pos_alt, n * s->pos_bytes)) ||
        (rc = dalloc((char **)&s->xref, n * s->pos_bytes)) ||
        (rc = dalloc((char **)&s->xwrap, n * s->pos_bytes)) ||
        (rc = dalloc((char **)&s->pos_spare, n * s->pos_bytes)))
        return bail(rc);
    for (int c = 0; c < 3 * ndim + 2; c++)
        if ((rc = dalloc(&s->spare_d[c], n))) return bail(rc);
    for (int c = 0; c < kPermI; c++)
        if ((rc = dalloc(&s->spare_i[c], n))) return bail(rc);
    for (int k = 0; k < 3; k++) {
        if ((rc = dalloc(&s->v[k], n)) || (rc = dalloc(&s->a[k], n)) ||
            (rc = dalloc(&s->f[k], n)))
            return bail(rc);
    }
    if ((rc = dalloc(&s->m, n)) || (rc = dalloc(&s->diam, n)) ||
        (rc = dalloc(&s->type, n)) || (rc = dalloc(&s->orig, n)) ||
        (rc = dalloc(&s->rank, n)) || (rc = dalloc(&s->stage, n * ndim)) ||
        (rc = dalloc(&s->stage_small, n)) || (rc = dalloc(&s->keys, n)) ||
        (rc = dalloc(&s->tkeys, n)) || (rc = dalloc(&s->vals, n)) ||
        (rc = dalloc(&s->tvals, n)) ||
        (rc = dalloc(&s->hist, (size_t)256 * (radix_blocks((int)n) + 1))) ||
        (rc = dalloc(&s->stats, 2)) ||
        (rc = dalloc(&s->bpart, 2 * (size_t)build_blocks((int)n))) ||
        (rc = dalloc(&s->part, (size_t)kRedWidth * force_blocks((int)n, 8))) ||
        (rc = dalloc(&s->out, kRedWidth)) || (rc = dalloc(&s->fcnt, 1)) ||
        (rc = dalloc(&s->fgpart, 3 * kFinBlocks)) ||
        (rc = dalloc(&s->dflag, 2 * kFlagRing + 4)) || (rc = dalloc(&s->tinfo, 4)))
        return bail(rc);
    if (hipHostMalloc((void **)&s->hflag, (2 * kFlagRing + 4) * sizeof(uint32_t),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&s->hflag_dev, s->hflag, 0) != hipSuccess ||
        hipHostMalloc((void **)&s->hrep, sizeof(BuildReport),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&s->hrep_dev, s->hrep, 0) != hipSuccess ||
        hipHostMalloc((void **)&s->hout, kRedWidth * sizeof(double), 0) != hipSuccess ||
        hipHostMalloc((void **)&s->htinfo, 4 * sizeof(int), 0) != hipSuccess)
        return bail(fail(PARM_E_NOMEM, "pinned host allocation failed"));
    hipStream_t st = s->stream;
    const int nb = (int)((n + 255) / 256);
    // zero state; orig = identity; no list members; type 0
    for (int k = 0; k < 3; k++) {
        if (hipMemsetAsync(s->v[k], 0, n * sizeof(double), st) != hipSuccess ||
            hipMemsetAsync(s->a[k], 0, n * sizeof(double), st) != hipSuccess ||
            hipMemsetAsync(s->f[k], 0, n * sizeof(double), st) != hipSuccess)
            return bail(fail(PARM_E_DEVICE, "memset failed"));
    }
    (void)hipMemsetAsync(s->pos, 0, n * s->pos_bytes, st);
    (void)hipMemsetAsync(s->pos_alt, 0, n * s->pos_bytes, st);
    (void)hipMemsetAsync(s->xref, 0, n * s->pos_bytes, st);
    (void)hipMemsetAsync(s->type, 0, n * sizeof(int32_t), st);
    (void)hipMemsetAsync(s->diam, 0, n * sizeof(double), st);
    (void)hipMemsetAsync(s->dflag, 0, (2 * kFlagRing + 4) * sizeof(uint32_t), st);
    (void)hipMemsetAsync(s->fcnt, 0, sizeof(uint32_t), st);
    for (int k = 0; k < 2 * kFlagRing + 4; k++) s->hflag[k] = 0;
    hipLaunchKernelGGL(k_iota, dim3(nb), dim3(256), 0, st, (int)n, s->orig);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, (int)n, s->rank, -1);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return bail(fail(PARM_E_DEVICE, "initialisation kernels failed"));
    if (s->n > 0) {
        rc = upload_scalar<double>(s, masses, s->m);
        if (rc) return bail(rc);
    }
    *out = s;
    return PARM_OK;
}

int parm_system_destroy(parm_system *sys) {
    destroy(sys);
    return PARM_OK;
}

int parm_system_info(const parm_system *s, int *ndim, int *precision, int64_t *natoms) {
    PARM_CHECK(check_sys(s));
    if (ndim) *ndim = s->ndim;
    if (precision) *precision = s->prec;
    if (natoms) *natoms = s->n;
    return PARM_OK;
}

int parm_set_box(parm_system *s, const double *L) {
    PARM_CHECK(check_sys(s));
    if (!L) return fail(PARM_E_INVALID, "L is null");
    return apply_box(s, L);
}

int parm_get_box(parm_system *s, double *L) {
    PARM_CHECK(check_sys(s));
    if (!L) return fail(PARM_E_INVALID, "L is null");
    for (int k = 0; k < s->ndim; k++) L[k] = s->box.L[k];
    return PARM_OK;
}

int parm_set_state(parm_system *s, const double *x, const double *v, const double *a,
                   const double *f) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    if (x) PARM_CHECK(upload_vec(s, 0, x));
    if (v) PARM_CHECK(upload_vec(s, 1, v));
    if (a) PARM_CHECK(upload_vec(s, 2, a));
    if (f) PARM_CHECK(upload_vec(s, 3, f));
    return PARM_OK;
}

int parm_get_state(parm_system *s, double *x, double *v, double *a, double *f) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    if (x) PARM_CHECK(download_vec(s, 0, x));
    if (v) PARM_CHECK(download_vec(s, 1, v));
    if (a) PARM_CHECK(download_vec(s, 2, a));
    if (f) PARM_CHECK(download_vec(s, 3, f));
    return PARM_OK;
}

int parm_set_masses(parm_system *s, const double *masses) {
    PARM_CHECK(check_sys(s));
    if (!masses && s->n > 0) return fail(PARM_E_INVALID, "masses is null");
    PARM_CHECK(set_device(s));
    return upload_scalar<double>(s, masses, s->m);
}

int parm_nlist_set(parm_system *s, double skin, int64_t nmembers, const int32_t *member,
                   const double *diameter) {
    PARM_CHECK(check_sys(s));
    if (!(skin >= 0) || !std::isfinite(skin))
        return fail(PARM_E_INVALID, "skin must be finite and >= 0");
    if (nmembers < 0 || nmembers > s->n)
        return fail(PARM_E_INVALID, "nmembers out of range");
    if (nmembers > 0 && (!member || !diameter))
        return fail(PARM_E_INVALID, "null member/diameter array");
    std::vector<int32_t> rank(s->n, -1);
    std::vector<double> dia(s->n, 0.0);
    double dmax = 0.0, dmin = 0.0;
    for (int64_t k = 0; k < nmembers; k++) {
        int32_t a = member[k];
        if (a < 0 || a >= s->n)
            return fail(PARM_E_INVALID, "member index out of range");
        if (rank[a] >= 0)
            return fail(PARM_E_INVALID,
                        "Cannot add AtomID to SubGroup: it already exists.");
        if (!(diameter[k] >= 0) || !std::isfinite(diameter[k]))
            return fail(PARM_E_INVALID, "diameters must be finite and >= 0");
        rank[a] = (int32_t)k;
        dia[a] = diameter[k];
        dmax = std::max(dmax, diameter[k]);
        dmin = k == 0 ? diameter[k] : std::min(dmin, diameter[k]);
    }
    PARM_CHECK(set_device(s));
    if (s->n > 0) {
        PARM_CHECK(upload_scalar<int32_t>(s, rank.data(), s->rank));
        PARM_CHECK(upload_scalar<double>(s, dia.data(), s->diam));
    }
    s->skin = skin;
    s->dmax = dmax;
    s->dmin = dmin;
    s->nmem = nmembers;
    s->list_cfg = true;
    s->ignorechanged = true;
    // a new NeighborList has an empty PairList
    s->excl.clear();
    s->excl_dirty = true;
    s->built = false;
    s->npairs = 0;
    free_list(s);
    return PARM_OK;
}

int parm_nlist_update(parm_system *s, int force, int *rebuilt) {
    PARM_CHECK(check_sys(s));
    if (!s->box_set) return fail(PARM_E_STATE, "box not set (parm_set_box)");
    PARM_CHECK(set_device(s));
    return update_list(s, force != 0, rebuilt);
}

int parm_nlist_ignore(parm_system *s, int64_t npairs, const int32_t *a, const int32_t *b) {
    PARM_CHECK(check_sys(s));
    if (!s->list_cfg) return fail(PARM_E_STATE, "neighbour list not set (parm_nlist_set)");
    if (npairs < 0 || (npairs > 0 && (!a || !b)))
        return fail(PARM_E_INVALID, "bad exclusion pair arrays");
    std::vector<uint64_t> add;
    add.reserve((size_t)npairs);
    for (int64_t k = 0; k < npairs; k++) {
        if (a[k] < 0 || a[k] >= s->n || b[k] < 0 || b[k] >= s->n)
            return fail(PARM_E_INVALID, "exclusion atom index out of range");
        const uint32_t hi = (uint32_t)std::max(a[k], b[k]), lo = (uint32_t)std::min(a[k], b[k]);
        add.push_back(((uint64_t)hi << 32) | lo);
    }
    std::sort(add.begin(), add.end());
    std::vector<uint64_t> merged;
    merged.reserve(s->excl.size() + add.size());
    std::merge(s->excl.begin(), s->excl.end(), add.begin(), add.end(),
               std::back_inserter(merged));
    merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
    s->excl.swap(merged);
    s->excl_dirty = true;
    // NeighborList::ignore sets ignorechanged: the next update rebuilds
    s->ignorechanged = true;
    return PARM_OK;
}

int parm_nlist_ignore_size(parm_system *s, int64_t *n) {
    PARM_CHECK(check_sys(s));
    if (n) *n = (int64_t)s->excl.size();
    return PARM_OK;
}

int parm_nlist_which(parm_system *s, uint32_t *which) {
    PARM_CHECK(check_sys(s));
    if (which) *which = s->which;
    return PARM_OK;
}

int parm_nlist_numpairs(parm_system *s, int64_t *npairs) {
    PARM_CHECK(check_sys(s));
    if (npairs) *npairs = s->built ? s->npairs : 0;
    return PARM_OK;
}

int parm_nlist_pairs(parm_system *s, int64_t cap, int32_t *first, int32_t *last,
                     int64_t *npairs) {
    PARM_CHECK(check_sys(s));
    if (cap > 0 && (!first || !last)) return fail(PARM_E_INVALID, "null output array");
    PARM_CHECK(set_device(s));
    const int64_t np = s->built ? s->npairs : 0;
    if (npairs) *npairs = np;
    if (np == 0 || cap <= 0) return PARM_OK;
    // Export: read the full list back and keep, per member i, the entries of
    // lower member rank, sorted by rank -- the reference's pair order.
    const int n = s->n;
    std::vector<int32_t> orig(n), rank(n);
    std::vector<uint32_t> nb((size_t)s->cap * s->npad);
    const uint32_t *src = s->nbr;
    uint32_t *conv = nullptr;
    if (s->tiled) {
        // staged slots -> atom entries, in a temporary untiled copy
        PARM_CHECK(dalloc(&conv, nb.size()));
        hipError_t e = launch_tiled_to_global(s->ndim, s->ntiles, s->tiles, s->keys,
                                              s->cstart, s->grid,
                                              reinterpret_cast<const uint16_t *>(s->nbr),
                                              s->lcount, s->type, s->type_bits, s->tpa,
                                              s->npad, conv, s->stream);
        if (e != hipSuccess) {
            dfree(conv);
            return fail(PARM_E_DEVICE, std::string("pair export: ") + hipGetErrorString(e));
        }
        src = conv;
    }
    std::vector<uint16_t> lcnt((size_t)n * s->tpa);
    PARM_HIP(hipMemcpyAsync(lcnt.data(), s->lcount, sizeof(uint16_t) * lcnt.size(),
                            hipMemcpyDeviceToHost, s->stream));
    PARM_HIP(hipMemcpyAsync(orig.data(), s->orig, sizeof(int32_t) * n,
                            hipMemcpyDeviceToHost, s->stream));
    PARM_HIP(hipMemcpyAsync(rank.data(), s->rank, sizeof(int32_t) * n,
                            hipMemcpyDeviceToHost, s->stream));
    hipError_t ec = hipMemcpyAsync(nb.data(), src, sizeof(uint32_t) * nb.size(),
                                   hipMemcpyDeviceToHost, s->stream);
    if (ec == hipSuccess) ec = hipStreamSynchronize(s->stream);
    dfree(conv);
    if (ec != hipSuccess)
        return fail(PARM_E_DEVICE, std::string("pair export: ") + hipGetErrorString(ec));
    std::vector<int32_t> slot_of_rank(s->nmem, -1);
    for (int i = 0; i < n; i++)
        if (rank[i] >= 0) slot_of_rank[rank[i]] = i;
    int64_t w = 0;
    std::vector<int32_t> lower;
    for (int64_t r = 0; r < s->nmem; r++) {
        const int i = slot_of_rank[r];
        lower.clear();
        for (int sub = 0; sub < s->tpa; sub++) {
            const int nl = lcnt[(size_t)i * s->tpa + sub];
            for (int m = 0; m < nl; m++) {
                const size_t idx = ((size_t)m * s->npad + i) * s->tpa + sub;
                const int j = (int)(nb[idx] >> s->type_bits);
                if (rank[j] < r) lower.push_back(rank[j]);
            }
        }
        std::sort(lower.begin(), lower.end());
        for (int32_t rj : lower) {
            if (w < cap) {
                first[w] = orig[i];
                last[w] = orig[slot_of_rank[rj]];
            }
            w++;
        }
    }
    if (w != np)
        return fail(PARM_E_DEVICE, "pair export count mismatch (list not symmetric)");
    return PARM_OK;
}

int parm_pair_lj_attract_repulse(parm_system *s, int ntypes, const double *eps_rows,
                                 const double *sigma, const double *sigcut,
                                 const int32_t *atom_type) {
    PARM_CHECK(check_sys(s));
    return build_pair_table(s, kFormLJAttractRepulse, ntypes, eps_rows, sigma, sigcut,
                            atom_type);
}

int parm_pair_repulsion(parm_system *s, int ntypes, const double *eps,
                        const double *sigma, const double *exponent,
                        const int32_t *atom_type) {
    PARM_CHECK(check_sys(s));
    return build_pair_table(s, kFormRepulsion, ntypes, eps, sigma, exponent, atom_type);
}

int parm_pair_lj_cut(parm_system *s, int ntypes, const double *eps, const double *sigma,
                     const double *sigcut, const int32_t *atom_type) {
    PARM_CHECK(check_sys(s));
    return build_pair_table(s, kFormLJCut, ntypes, eps, sigma, sigcut, atom_type);
}

int parm_pair_lj_cut_indexed(parm_system *s, int ntypes, const double *eps_rows,
                             const double *sigma_rows, const double *sigcut,
                             const int32_t *atom_type) {
    PARM_CHECK(check_sys(s));
    return build_pair_table(s, kFormLJCutIndexed, ntypes, eps_rows, sigma_rows, sigcut,
                            atom_type);
}

int parm_pair_lj_repulse(parm_system *s, int ntypes, const double *eps, const double *sigma,
                         const int32_t *atom_type) {
    PARM_CHECK(check_sys(s));
    return build_pair_table(s, kFormLJRepulse, ntypes, eps, sigma, nullptr, atom_type);
}

int parm_pair_repulsion_indexed(parm_system *s, int ntypes, const double *eps_rows,
                                const double *sigma_rows, const double *exponent,
                                const int32_t *atom_type) {
    PARM_CHECK(check_sys(s));
    return build_pair_table(s, kFormRepulsionIndexed, ntypes, eps_rows, sigma_rows, exponent,
                            atom_type);
}

int parm_set_forces(parm_system *s, int and_a) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(require_ready(s));
    PARM_CHECK(set_device(s));
    return set_forces(s, and_a != 0);
}

int parm_add_pair_forces(parm_system *s) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(require_ready(s));
    PARM_CHECK(set_device(s));
    StepCoef c{};
    PARM_HIP(launch_force(s->ndim, s->prec, kStageForceAdd, false, s->arrays(),
                          s->listview(), s->pairview(), s->box, c, s->redbuf(kSlotScratch),
                          s->stream));
    return PARM_OK;
}

int parm_pair_sums(parm_system *s, double *U, double *W, double *S, uint64_t *contacts,
                   uint64_t *overlaps) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(require_ready(s));
    PARM_CHECK(set_device(s));
    StepCoef c{};
    PARM_HIP(launch_force(s->ndim, s->prec, kStageSums, true, s->arrays(), s->listview(),
                          s->pairview(), s->box, c, s->redbuf(kSlotScratch), s->stream));
    PARM_CHECK(read_out(s));
    const double *o = s->hout;
    if (U) *U = o[0];
    if (W) *W = o[1];
    if (S)
        for (int k = 0; k < s->ndim * s->ndim; k++) S[k] = o[2 + k];
    if (contacts) *contacts = (uint64_t)std::llround(o[11] / 2);
    if (overlaps) *overlaps = (uint64_t)std::llround(o[12] / 2);
    return PARM_OK;
}

}  // extern "C"

// ------------------------------------------- sums without allocation
// Scratch of parm_pair_sums_noalloc: the per-workgroup partials of one force
// launch (kRedWidth fields x force_blocks(n, 8) workgroups) and the reduced
// fields, for systems of at most kNoAllocMaxAtoms atoms.  Each GPU holds its
// own copy for the life of the process; host code reaches it only through
// the per-device address from hipGetSymbolAddress.
constexpr int kNoAllocMaxAtoms = 1 << 21;
constexpr size_t kNoAllocParts = (size_t)kNoAllocMaxAtoms * 8 / kBlock;  // workgroups
static __device__ __align__(256) double scratch[kRedWidth * kNoAllocParts + kRedWidth];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The calling thread's current device's scratch address (looked up once per
// device, with that device current).
int scratch_address(int device, double **out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(PARM_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PARM_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                           hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return PARM_OK;
}

// Reduced fields -> the caller's layout: U, W, S[9] (ndim x ndim used),
// contacts, overlaps (each pair counted once).
// (red == nullptr: all zero, for a system with no atoms).
__global__ void k_sums_out(const double *__restrict__ red, double *__restrict__ out) {
    const int k = threadIdx.x;
    if (k >= 13) return;
    const double v = red ? red[k] : 0.0;
    out[k] = k < 11 ? v : v / 2;
}
}  // namespace

extern "C" {

int parm_pair_sums_noalloc(parm_system *s, double *out) {
    PARM_CHECK(check_sys(s));
    if (!out) return fail(PARM_E_INVALID, "out is null");
    if (s->n > kNoAllocMaxAtoms)
        return fail(PARM_E_UNSUPPORTED,
                    "parm_pair_sums_noalloc: at most " + std::to_string(kNoAllocMaxAtoms) +
                        " atoms (its scratch is a fixed device array)");
    PARM_CHECK(require_ready(s));
    PARM_CHECK(set_device(s));
    double *scr = nullptr;
    PARM_CHECK(scratch_address(s->device, &scr));
    ReduceBuf R = s->redbuf(kSlotScratch);
    R.part = scr;
    R.pstride = force_blocks(std::max(s->n, 1), 8);  // <= kNoAllocParts
    R.out = scr + kRedWidth * kNoAllocParts;
    if (s->n > 0) {
        StepCoef c{};
        PARM_HIP(launch_force(s->ndim, s->prec, kStageSums, true, s->arrays(), s->listview(),
                              s->pairview(), s->box, c, R, s->stream));
        hipLaunchKernelGGL(k_sums_out, dim3(1), dim3(64), 0, s->stream, R.out, out);
    } else {
        // an empty system has no pairs: zero sums, by a kernel (no memset)
        hipLaunchKernelGGL(k_sums_out, dim3(1), dim3(64), 0, s->stream,
                           (const double *)nullptr, out);
    }
    PARM_HIP(hipGetLastError());
    return PARM_OK;
}

int parm_kinetic_energy(parm_system *s, const double *originv, double *K) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    ReduceBuf R = s->redbuf(kSlotScratch);
    R.skin = INFINITY;
    PARM_HIP(launch_atom_sums(s->ndim, 0, s->arrays(), originv, R, s->stream));
    PARM_CHECK(read_out(s));
    if (K) *K = s->hout[0];
    return PARM_OK;
}

int parm_momentum(parm_system *s, double *p, double *mass) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    ReduceBuf R = s->redbuf(kSlotScratch);
    R.skin = INFINITY;
    PARM_HIP(launch_atom_sums(s->ndim, 0, s->arrays(), nullptr, R, s->stream));
    PARM_CHECK(read_out(s));
    if (p)
        for (int k = 0; k < s->ndim; k++) p[k] = s->hout[1 + k];
    if (mass) *mass = s->hout[4];
    return PARM_OK;
}

int parm_force_stats(parm_system *s, double *fmax, double *fmean) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    ReduceBuf R = s->redbuf(kSlotScratch);
    R.skin = INFINITY;
    PARM_HIP(launch_atom_sums(s->ndim, 1, s->arrays(), nullptr, R, s->stream));
    PARM_CHECK(read_out(s));
    if (fmax) *fmax = s->hout[13];
    if (fmean) *fmean = s->n > 0 ? s->hout[0] / s->n : 0.0;
    return PARM_OK;
}

int parm_initialize(parm_system *s) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(require_ready(s));
    PARM_CHECK(set_device(s));
    PARM_CHECK(update_list(s, false, nullptr));
    PARM_CHECK(set_forces(s, true));
    PARM_CHECK(update_list(s, false, nullptr));
    return PARM_OK;
}

int parm_verlet_steps(parm_system *s, double dt, int64_t nsteps) {
    PARM_CHECK(check_sys(s));
    if (nsteps < 0) return fail(PARM_E_INVALID, "nsteps < 0");
    PARM_CHECK(set_device(s));
    StepCoef c{};
    c.k0 = dt;
    c.k1 = dt * dt / 2;
    c.k2 = dt / 2;
    return run_steps(s, false, c, nsteps);
}

int parm_damped_steps(parm_system *s, double dt, double damping, int64_t nsteps) {
    PARM_CHECK(check_sys(s));
    if (dt <= 0) return fail(PARM_E_INVALID, "Collection::CollectionSol: dt >= 0");
    if (nsteps < 0) return fail(PARM_E_INVALID, "nsteps < 0");
    PARM_CHECK(set_device(s));
    // CollectionDamped::set_constants (src/collection.cpp:342-354)
    double c0, c1, c2;
    if (damping <= 0.0) {
        c0 = 1;
        c1 = 1;
        c2 = .5;
    } else {
        double dampdt = damping * dt;
        c0 = std::exp(-dampdt);
        c1 = (-std::expm1(-dampdt)) / dampdt;
        c2 = (1 - c1) / dampdt;
    }
    StepCoef c{};
    c.k0 = c1 * dt;
    c.k1 = c2 * dt * dt;
    c.k2 = dt * (c1 - c2);
    c.k3 = c0;
    c.k4 = dt * c2;
    return run_steps(s, true, c, nsteps);
}

int parm_autotune(parm_system *s, double dt, int damped) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(require_ready(s));
    if (!(dt > 0) || !std::isfinite(dt)) return fail(PARM_E_INVALID, "dt must be finite and > 0");
    PARM_CHECK(set_device(s));
    if (!s->built) PARM_CHECK(update_list(s, false, nullptr));
    // the run's own fused kernel; its outputs go to scratch, so the
    // coefficients only need to be finite
    StepCoef c{};
    c.k0 = dt;
    c.k1 = dt * dt / 2;
    c.k2 = dt / 2;
    c.k3 = 1.0;
    c.k4 = dt / 2;
    PARM_CHECK(autotune(s, damped ? kStageDampedKickDrift : kStageVerletKickDrift, c));
    PARM_HIP(hipStreamSynchronize(s->stream));
    return PARM_OK;
}

int parm_profile_enable(parm_system *s, int enable) {
    PARM_CHECK(check_sys(s));
    if (enable < 0) return fail(PARM_E_INVALID, "profile stride must be >= 0");
    s->prof = enable != 0;
    s->prof_every = std::max(enable, 1);
    s->prof_seq = 0;
    return PARM_OK;
}

int parm_profile_read(parm_system *s, int reset, double *step_ms, int64_t *step_launches,
                      double *rebuild_ms, int64_t *rebuilds) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    PARM_HIP(hipStreamSynchronize(s->stream));
    auto total = [](const std::vector<std::pair<hipEvent_t, hipEvent_t>> &v, double *ms) {
        double t = 0.0;
        for (const auto &p : v) {
            float e = 0.0f;
            if (hipEventElapsedTime(&e, p.first, p.second) == hipSuccess) t += e;
        }
        if (ms) *ms = t;
    };
    total(s->prof_step, step_ms);
    total(s->prof_rebuild, rebuild_ms);
    if (step_launches) *step_launches = (int64_t)s->prof_step.size();
    if (rebuilds) *rebuilds = (int64_t)s->prof_rebuild.size();
    if (reset) {
        s->prof_step.clear();
        s->prof_rebuild.clear();
        s->ev_used = 0;
    }
    return PARM_OK;
}

int parm_list_info(parm_system *s, int *tiled, int *lanes, int *capacity) {
    PARM_CHECK(check_sys(s));
    if (tiled) *tiled = s->tiled ? 1 : 0;
    if (lanes) *lanes = s->tpa;
    if (capacity) *capacity = s->cap;
    return PARM_OK;
}

int parm_synchronize(parm_system *s) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    PARM_HIP(hipStreamSynchronize(s->stream));
    return PARM_OK;
}

// ------------------------------------------------------------ snapshots
int parm_snapshot_begin(parm_system *s, int what, int64_t *ticket) {
    PARM_CHECK(check_sys(s));
    if (!ticket) return fail(PARM_E_INVALID, "ticket is null");
    if (what & ~3) return fail(PARM_E_INVALID, "what: bit 0 positions, bit 1 velocities");
    PARM_CHECK(set_device(s));
    const size_t nv = (size_t)s->n * s->ndim;
    parm_system::Snap &q = s->snap[s->snap_next % parm_system::kSnapRing];
    if (!s->out_stream &&
        hipStreamCreateWithFlags(&s->out_stream, hipStreamNonBlocking) != hipSuccess)
        return fail(PARM_E_DEVICE, "snapshot stream creation failed");
    if (!q.dev) {
        PARM_CHECK(dalloc(&q.dev, 2 * std::max<size_t>(nv, 1)));
        if (hipHostMalloc((void **)&q.host, sizeof(double) * 2 * std::max<size_t>(nv, 1), 0) !=
            hipSuccess)
            return fail(PARM_E_NOMEM, "pinned snapshot allocation failed");
        PARM_HIP(hipEventCreateWithFlags(&q.ready, hipEventDisableTiming));
        PARM_HIP(hipEventCreateWithFlags(&q.done, hipEventDisableTiming));
    } else if (q.ticket >= 0) {
        // the slot's previous copy must have left before it is refilled
        PARM_HIP(hipEventSynchronize(q.done));
    }
    hipStream_t st = s->stream;
    if (nv > 0) {
        if (what & 1) PARM_HIP(launch_download(s->ndim, 0, s->arrays(), s->orig, q.dev, st));
        if (what & 2)
            PARM_HIP(launch_download(s->ndim, 1, s->arrays(), s->orig, q.dev + nv, st));
    }
    PARM_HIP(hipEventRecord(q.ready, st));
    PARM_HIP(hipStreamWaitEvent(s->out_stream, q.ready, 0));
    if (nv > 0 && what)
        PARM_HIP(hipMemcpyAsync(q.host, q.dev, sizeof(double) * 2 * nv, hipMemcpyDeviceToHost,
                                s->out_stream));
    PARM_HIP(hipEventRecord(q.done, s->out_stream));
    for (int k = 0; k < 3; k++) q.L[k] = s->box.L[k];
    q.what = what;
    q.ticket = s->snap_next++;
    *ticket = q.ticket;
    return PARM_OK;
}

namespace {
// The ring slot of a snapshot still held, after its copy arrived.
int snapshot_slot(parm_system *s, int64_t ticket, parm_system::Snap **out) {
    if (ticket < 0 || ticket >= s->snap_next ||
        ticket < s->snap_next - parm_system::kSnapRing)
        return fail(PARM_E_INVALID, "snapshot " + std::to_string(ticket) +
                                        " is not held (only the last " +
                                        std::to_string(parm_system::kSnapRing) + " are)");
    parm_system::Snap &q = s->snap[ticket % parm_system::kSnapRing];
    if (q.ticket != ticket) return fail(PARM_E_STATE, "snapshot slot reused");
    PARM_HIP(hipEventSynchronize(q.done));
    *out = &q;
    return PARM_OK;
}
}  // namespace

int parm_snapshot_read(parm_system *s, int64_t ticket, double *x, double *v, double *L) {
    PARM_CHECK(check_sys(s));
    PARM_CHECK(set_device(s));
    parm_system::Snap *q = nullptr;
    PARM_CHECK(snapshot_slot(s, ticket, &q));
    const size_t nv = (size_t)s->n * s->ndim;
    if (x) {
        if (!(q->what & 1)) return fail(PARM_E_STATE, "snapshot holds no positions");
        std::memcpy(x, q->host, sizeof(double) * nv);
    }
    if (v) {
        if (!(q->what & 2)) return fail(PARM_E_STATE, "snapshot holds no velocities");
        std::memcpy(v, q->host + nv, sizeof(double) * nv);
    }
    if (L)
        for (int k = 0; k < s->ndim; k++) L[k] = q->L[k];
    return PARM_OK;
}

int parm_snapshot_write_xyz(parm_system *s, int64_t ticket, const char *path, int append,
                            const char *comment, const char *elements, int precision) {
    PARM_CHECK(check_sys(s));
    if (!path) return fail(PARM_E_INVALID, "path is null");
    if (precision < 1 || precision > 40) return fail(PARM_E_INVALID, "precision in [1, 40]");
    PARM_CHECK(set_device(s));
    parm_system::Snap *q = nullptr;
    PARM_CHECK(snapshot_slot(s, ticket, &q));
    if (!(q->what & 1)) return fail(PARM_E_STATE, "snapshot holds no positions");
    if (comment && std::strchr(comment, '\n'))
        return fail(PARM_E_INVALID, "the comment must be one line");
    // one frame: the atom count, a comment line, then per atom an element
    // letter and the ndim components of diff(0, x) = remainder(-x, L), as
    // the reference's drivers write them (src/bin/LJatoms.cpp:130-158,
    // src/bin/packer.cpp:150-185); %g at `precision` digits is what an
    // ostream with that precision prints
    const int d = s->ndim;
    std::string out;
    out.reserve((size_t)s->n * (size_t)(d * (precision + 9) + 3) + 64);
    char buf[96];
    std::snprintf(buf, sizeof buf, "%d\n", s->n);
    out += buf;
    if (comment) out += comment;
    out += '\n';
    for (int i = 0; i < s->n; i++) {
        out += elements ? elements[i] : 'C';
        for (int k = 0; k < d; k++) {
            const double r = std::remainder(0.0 - q->host[(size_t)i * d + k], q->L[k]);
            std::snprintf(buf, sizeof buf, "\t%.*g", precision, r);
            out += buf;
        }
        out += '\n';
    }
    FILE *fh = std::fopen(path, append ? "ab" : "wb");
    if (!fh) return fail(PARM_E_INVALID, std::string("cannot open ") + path);
    const size_t wrote = std::fwrite(out.data(), 1, out.size(), fh);
    const int closed = std::fclose(fh);
    if (wrote != out.size() || closed != 0)
        return fail(PARM_E_DEVICE, std::string("write to ") + path + " failed");
    return PARM_OK;
}

}  // extern "C"

// ================================================ nonlinear CG minimiser
// CollectionNLCG / CollectionNLCGFixedL (src/collection.hpp:400-474;
// src/collection.cpp:494-870) on one device system whose collection has the
// system's pair interaction and NeighborList only.  The scalar logic is the
// reference's, in its order; every O(N) loop is a device pass.
struct parm_nlcg {
    parm_system *s = nullptr;
    bool fixedL = false;
    // parameters
    double dt = 0, seceps = 0;
    unsigned secmax = 0;
    double kappa = 0, alphamax = 2.0, afrac = 0, dxmax = 100, stepmax = 1e-3, kmax = 0;
    double P0 = 0;
    // kept between iterations
    double Knew = 0, k = 0, vl = 0, fl = 0, al = 0;
    // tracking
    double alpha = 0, beta = 0, betaused = 0, dxsum = 0, alphavmax = 0, maxdV = 0;
    unsigned sec = 0;
};

namespace {

double box_V(const parm_system *s) {
    double V = 1.0;
    for (int k = 0; k < s->ndim; k++) V *= s->box.L[k];
    return V;
}

// get_length_squared (src/collection.cpp:627-635)
double length_squared(const parm_system *s) {
    return s->ndim == 3 ? std::pow(box_V(s), 2.0 / 3.0) : box_V(s);
}

// One vector pass (assignment op, then the sums), read back.
int nlcg_pass(parm_nlcg *c, int op, double beta, double lfac_all, double out[5]) {
    parm_system *s = c->s;
    const double lf[3] = {lfac_all, lfac_all, lfac_all};
    ReduceBuf R = s->redbuf(kSlotScratch);
    PARM_HIP(launch_nlcg_vec(s->ndim, op, s->arrays(), beta, lf, R, s->stream));
    if (!out) return PARM_OK;
    PARM_CHECK(read_out(s));
    for (int q = 0; q < 5; q++) out[q] = s->hout[q];
    return PARM_OK;
}

// stepx (src/collection.cpp:608-625; FixedL :855-870): x *= Lfac, x += v dx,
// then the box edges (all, or the first one) times Lfac.
int nlcg_stepx(parm_nlcg *c, double dx) {
    parm_system *s = c->s;
    const int d = s->ndim;
    const double Lfac = c->fixedL ? std::exp(dx * c->vl / c->kappa)
                                  : std::exp(dx * c->vl / (c->kappa * d));
    double lf[3] = {Lfac, Lfac, Lfac};
    if (c->fixedL) lf[1] = lf[2] = 1.0;
    PARM_HIP(launch_nlcg_stepx(d, s->arrays(), lf, dx, s->stream));
    double L[3];
    for (int k = 0; k < d; k++) L[k] = s->box.L[k];
    if (c->fixedL)
        L[0] *= Lfac;
    else
        for (int k = 0; k < d; k++) L[k] *= Lfac;
    return apply_box(s, L);
}

// Collection::set_forces_get_pressure(false) (src/collection.cpp:181-208):
// f = pair forces, a = 0 for fixed atoms; returns sum r.f.
int forces_get_pressure(parm_system *s, double *W) {
    PARM_CHECK(set_forces(s, false));
    double zero[3] = {1.0, 1.0, 1.0};
    ReduceBuf R = s->redbuf(kSlotScratch);
    PARM_HIP(launch_nlcg_vec(s->ndim, kNlcgFixedA0, s->arrays(), 0.0, zero, R, s->stream));
    StepCoef cf{};
    PARM_HIP(launch_force(s->ndim, s->prec, kStageSums, true, s->arrays(), s->listview(),
                          s->pairview(), s->box, cf, s->redbuf(kSlotScratch), s->stream));
    PARM_CHECK(read_out(s));
    *W = s->hout[1];
    return PARM_OK;
}

// CollectionNLCG::set_forces(constraints_and_a, setV) (src/collection.cpp:
// 537-569).
int nlcg_set_forces(parm_nlcg *c, bool constraints_and_a, bool setV) {
    parm_system *s = c->s;
    const double V = box_V(s);
    if (setV) {
        double interacP = 0;
        PARM_CHECK(forces_get_pressure(s, &interacP));
        c->fl = ((interacP / s->ndim) - (c->P0 * V)) / c->kappa;
        if (constraints_and_a) {
            c->al = c->fl;
            c->vl = c->fl;
        }
    } else {
        PARM_CHECK(set_forces(s, false));
    }
    if (constraints_and_a) {
        double o[5];
        PARM_CHECK(nlcg_pass(c, kNlcgAssignAV, 0.0, 1.0, o));
        c->Knew = o[kNlcgFA] / length_squared(s) + c->fl * c->al;  // fdota
    }
    return PARM_OK;
}

// fdotv / vdotv / fdota (src/collection.cpp:650-683) from one pass.
int nlcg_dots(parm_nlcg *c, double *fdota, double *fdotv, double *vdotv) {
    double o[5];
    PARM_CHECK(nlcg_pass(c, kNlcgDots, 0.0, 1.0, o));
    const double l2 = length_squared(c->s);
    if (fdota) *fdota = o[kNlcgFA] / l2 + c->fl * c->al;
    if (fdotv) *fdotv = o[kNlcgFV] / l2 + c->fl * c->vl;
    if (vdotv) *vdotv = o[kNlcgVV] / l2 + c->vl * c->vl;
    return PARM_OK;
}

int nlcg_update_trackers(parm_nlcg *c) { return update_list(c->s, false, nullptr); }

// CollectionNLCG::timestep (src/collection.cpp:685-835).
int nlcg_timestep(parm_nlcg *c) {
    PARM_CHECK(nlcg_stepx(c, c->dt));
    PARM_CHECK(nlcg_set_forces(c, false, true));
    double fv = 0;
    PARM_CHECK(nlcg_dots(c, nullptr, &fv, nullptr));
    double eta0 = -fv;  // slope at x0 - dt
    double eta;
    PARM_CHECK(nlcg_stepx(c, -c->dt));
    PARM_CHECK(nlcg_update_trackers(c));
    c->alpha = -c->dt;
    PARM_CHECK(nlcg_set_forces(c, false, true));
    c->dxsum = 0;
    double vdv = 0;
    PARM_CHECK(nlcg_dots(c, nullptr, &fv, &vdv));
    const int d = c->s->ndim;
    // secant steps along v until the slope along v is small
    for (c->sec = 0; c->sec < c->secmax; c->sec++) {
        eta = -fv;  // the slope at x; fdotv of the last force evaluation
        double alphafac = -eta / std::fabs(eta0 - eta);
        if (std::fabs(eta0 - eta) <= 1e-12 * std::fabs(eta)) {
            alphafac = c->alphamax > 0 ? c->alphamax : 1.1;
            c->sec = c->sec > 0 ? c->sec * 2 - 1 : 1;
        }
        if (c->alphamax > 0 && alphafac > c->alphamax) alphafac = c->alphamax;
        if (c->alphamax > 0 && alphafac < -c->alphamax) alphafac = -c->alphamax;
        c->alpha = std::fabs(c->alpha) * alphafac;
        const double newdxsum = std::fabs(c->dxsum + c->alpha);
        if (c->dxmax > 0 && newdxsum > c->dxmax) {
            c->k = 0;
            break;
        }
        const double dVoverV =
            std::expm1(std::fabs(c->dxsum + c->alpha) * c->vl / (c->kappa * d));
        if (c->maxdV > 0 && dVoverV > c->maxdV) {
            c->k = 0;
            break;
        }
        c->dxsum += c->alpha;
        PARM_CHECK(nlcg_stepx(c, c->alpha));
        PARM_CHECK(nlcg_set_forces(c, false, true));
        eta0 = eta;
        if (c->alpha * c->alpha * vdv < c->seceps * c->seceps) break;
        if ((c->sec > 1) && (c->afrac > 0) && (std::fabs(c->alpha) < std::fabs(c->dxsum)) &&
            (std::fabs(c->alpha) / std::fabs(c->dxsum) < c->afrac))
            break;
        if (c->stepmax > 0 && c->dxsum * c->dxsum * vdv > c->stepmax * c->stepmax) {
            c->k = 0;
            break;
        }
        // the next iteration's slope (the reference's fdotv at its top)
        PARM_CHECK(nlcg_dots(c, nullptr, &fv, nullptr));
    }
    c->alphavmax = std::sqrt(c->alpha * c->alpha * vdv);
    const double Kold = c->Knew;
    double Kmid = 0;
    PARM_CHECK(nlcg_dots(c, &Kmid, nullptr, nullptr));
    // a = f, al = fl, Knew = fdota
    double o[5];
    PARM_CHECK(nlcg_pass(c, kNlcgAssignA, 0.0, 1.0, o));
    c->al = c->fl;
    c->Knew = o[kNlcgFA] / length_squared(c->s) + c->fl * c->al;
    c->beta = (c->Knew - Kmid) / Kold;
    c->betaused = c->beta;
    c->k++;
    if (c->k >= c->kmax || std::isinf(c->betaused) || std::isnan(c->betaused) ||
        c->betaused <= 0) {
        c->k = 0;
        c->betaused = 0;
    } else if (c->betaused > 1) {
        c->betaused = 1;
    }
    PARM_CHECK(nlcg_pass(c, kNlcgAssignV, c->betaused, 1.0, nullptr));
    c->vl = c->al + c->betaused * c->vl;
    return PARM_OK;
}

int check_nlcg(const parm_nlcg *c) {
    if (!c || !c->s) return fail(PARM_E_INVALID, "null parm_nlcg");
    return PARM_OK;
}

}  // namespace

extern "C" {

int parm_nlcg_create(parm_system *s, double dt, double P0, double kappa, double kmax,
                     int secmax, double seceps, int fixed_L, parm_nlcg **out) {
    PARM_CHECK(check_sys(s));
    if (!out) return fail(PARM_E_INVALID, "out is null");
    *out = nullptr;
    if (!(kappa > 0)) return fail(PARM_E_INVALID, "kappa must be > 0");
    if (secmax < 0) return fail(PARM_E_INVALID, "secmax must be >= 0");
    PARM_CHECK(require_ready(s));
    PARM_CHECK(set_device(s));
    // Collection's constructor runs initialize() with Collection::set_forces
    // (src/collection.cpp:3-19): the NLCG members start at zero
    PARM_CHECK(parm_initialize(s));
    parm_nlcg *c = new parm_nlcg();
    c->s = s;
    c->fixedL = fixed_L != 0;
    c->dt = dt;
    c->P0 = P0;
    c->kappa = kappa;
    c->kmax = kmax;
    c->secmax = (unsigned)secmax;
    c->seceps = seceps;
    *out = c;
    return PARM_OK;
}

int parm_nlcg_destroy(parm_nlcg *c) {
    delete c;
    return PARM_OK;
}

int parm_nlcg_set_limits(parm_nlcg *c, double alphamax, double afrac, double dxmax,
                         double stepmax, double maxdV) {
    PARM_CHECK(check_nlcg(c));
    c->alphamax = alphamax;
    c->afrac = afrac;
    c->dxmax = dxmax;
    c->stepmax = stepmax;
    c->maxdV = maxdV;
    return PARM_OK;
}

int parm_nlcg_reset(parm_nlcg *c) {
    PARM_CHECK(check_nlcg(c));
    PARM_CHECK(set_device(c->s));
    // reset (src/collection.cpp:528-535): k = 0; set_forces(true, true)
    // (a = v = f); v = a; vl = al
    c->k = 0;
    PARM_CHECK(nlcg_set_forces(c, true, true));
    c->vl = c->al;
    return PARM_OK;
}

int parm_nlcg_set_params(parm_nlcg *c, double dt, double P0, double kappa, double kmax,
                         int secmax, double seceps) {
    PARM_CHECK(check_nlcg(c));
    if (!(kappa > 0)) return fail(PARM_E_INVALID, "kappa must be > 0");
    if (secmax < 0) return fail(PARM_E_INVALID, "secmax must be >= 0");
    c->dt = dt;
    c->P0 = P0;
    c->kappa = kappa;
    c->kmax = kmax;
    c->secmax = (unsigned)secmax;
    c->seceps = seceps;
    return PARM_OK;
}

int parm_nlcg_set_forces(parm_nlcg *c, int constraints_and_a, int setV) {
    PARM_CHECK(check_nlcg(c));
    PARM_CHECK(set_device(c->s));
    return nlcg_set_forces(c, constraints_and_a != 0, setV != 0);
}

int parm_nlcg_timestep(parm_nlcg *c, int64_t nsteps) {
    PARM_CHECK(check_nlcg(c));
    if (nsteps < 0) return fail(PARM_E_INVALID, "nsteps < 0");
    PARM_CHECK(set_device(c->s));
    for (int64_t q = 0; q < nsteps; q++) PARM_CHECK(nlcg_timestep(c));
    return PARM_OK;
}

int parm_nlcg_descend(parm_nlcg *c) {
    PARM_CHECK(check_nlcg(c));
    PARM_CHECK(set_device(c->s));
    // descend (src/collection.cpp:837-853)
    PARM_CHECK(nlcg_set_forces(c, false, true));
    PARM_CHECK(nlcg_pass(c, kNlcgDescend, 0.0, 1.0, nullptr));
    c->al = c->fl;
    c->vl = c->fl;
    PARM_CHECK(nlcg_stepx(c, c->dt));
    return nlcg_update_trackers(c);
}

int parm_nlcg_observables(parm_nlcg *c, double *H, double *P, double *K) {
    PARM_CHECK(check_nlcg(c));
    parm_system *s = c->s;
    PARM_CHECK(set_device(s));
    const double V = box_V(s);
    if (H || P) {
        double U = 0, W = 0;
        PARM_CHECK(parm_pair_sums(s, &U, &W, nullptr, nullptr, nullptr));
        if (H) *H = U + c->P0 * V;                 // hamiltonian (:571-573)
        if (P) *P = W / V / (double)s->ndim;       // pressure (:597-606)
    }
    if (K) {  // kinetic_energy, masses ignored (:575-590)
        const double Lfac = std::exp(c->vl / (c->kappa * s->ndim));
        double o[5];
        PARM_CHECK(nlcg_pass(c, kNlcgDots, 0.0, Lfac, o));
        *K = o[kNlcgKin] / 2.0;
    }
    return PARM_OK;
}

int parm_nlcg_info(parm_nlcg *c, double *out) {
    PARM_CHECK(check_nlcg(c));
    if (!out) return fail(PARM_E_INVALID, "out is null");
    const double v[12] = {c->alpha, c->beta,     c->betaused, c->dxsum,
                          c->alphavmax, (double)c->sec, c->k, c->vl,
                          c->fl,    c->al,       c->Knew,     c->maxdV};
    for (int q = 0; q < 12; q++) out[q] = v[q];
    return PARM_OK;
}

}  // extern "C"
