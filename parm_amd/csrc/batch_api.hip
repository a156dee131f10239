// batch_api.hip -- C ABI of the batched path (include/parm_hip.h, parm_batch_*).
//
// A parm_batch holds nsys independent systems, each equivalent to one ParM
// AtomVec + OriginBox + NeighborList + NListed pair + CollectionDamped /
// CollectionVerlet, with its own box and atom count.  The step loop issues no
// host-side decision: per-system skin tests and rebuilds run on the device,
// so a run of steps is captured once in a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/parm_hip.h"
#include "common.hpp"
#include "kernels.hpp"
#include "pairtable.hpp"

using namespace parm;

namespace {

constexpr double kPi = 3.14159265358979323846;

int bfail(int code, const std::string &msg) { return set_error(code, msg); }

#define PB_HIP(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return bfail(PARM_E_DEVICE, std::string("HIP error ") +            \
                                            hipGetErrorString(e_) + " in " +   \
                                            #call);                            \
    } while (0)
#define PB_CHECK(expr)                                                         \
    do {                                                                       \
        int r_ = (expr);                                                       \
        if (r_ != PARM_OK) return r_;                                          \
    } while (0)

template <typename T>
int balloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, sizeof(T) * count) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return bfail(PARM_E_NOMEM, "device allocation of " +
                                       std::to_string(sizeof(T) * count) + " bytes failed");
    }
    return PARM_OK;
}
template <typename T>
void bfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

struct parm_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int ndim = 3, prec = kFP64;
    int nsys = 0, ntot = 0, npad = 64;
    size_t pos_bytes = 32;
    double skin = 0.0;
    std::vector<int> off;           // nsys + 1
    std::vector<double> L;          // nsys * ndim
    std::vector<BatchGeom> geom_h;
    BatchGeom *geom = nullptr;

    void *pos = nullptr, *pos_alt = nullptr, *xref = nullptr;
    double *v[3] = {}, *a[3] = {}, *f[3] = {};
    double *m = nullptr, *diam = nullptr, *disp = nullptr, *stage = nullptr;
    int32_t *type = nullptr, *sys = nullptr;

    int cap = 0, tpa = 4, type_bits = 0, total_cells = 0;
    bool tuned = false, tpa_fixed = false, built = false;
    uint32_t *nbr = nullptr;
    int32_t *ncount = nullptr, *keys = nullptr, *cellidx = nullptr, *cstart = nullptr;
    void *xs = nullptr;        // positions in each system's cell order (rebuild scratch)
    double *dsort = nullptr;   // list diameters in the same order
    uint32_t *flags = nullptr, *which = nullptr, *gflag = nullptr, *hgflag = nullptr;
    int32_t *work = nullptr;   // worklist of systems to rebuild (count first)
    int32_t *wslice = nullptr; // first build slice of each listed system, then the total
    int sort_grid = 1;         // workgroups of the rebuild's sort launch
    int build_grid = 1;        // workgroups of the rebuild's list-build launch
    unsigned long long *nent = nullptr;  // list entries of each system (twice its pairs)

    int pair_kind = kPairNone, ntypes = 0, single = 1, harmonic = 0;
    PairParam *tab = nullptr;
    PairParam tab0{};  // host copy of tab[0]
    double *eatom = nullptr, *sysout = nullptr;  // 4 per atom; 4 x nsys sums

    hipGraphExec_t gexec = nullptr;
    // instrumentation: while on, runs launch eagerly and every fused step's
    // force launch and rebuild launch is bracketed by events
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_force, prof_rebuild;
    bool g_damped = false;
    StepCoef g_coef{};
    void *g_pos = nullptr;
    int g_tpa = 0;
    const uint32_t *g_nbr = nullptr;

    BatchArrays arrays(bool alt_out = false) const {
        BatchArrays A{};
        A.ntot = ntot;
        A.pos_in = pos;
        A.pos_out = alt_out ? pos_alt : pos;
        for (int k = 0; k < 3; k++) {
            A.v[k] = v[k];
            A.a[k] = a[k];
            A.f[k] = f[k];
        }
        A.m = m;
        A.type = type;
        A.sys = sys;
        A.diam = diam;
        A.xref = xref;
        A.xref_w = xref;
        A.disp = disp;
        A.geom = geom;
        return A;
    }
    ListView listview() const {
        ListView Lv{};
        Lv.nbr = nbr;
        Lv.ncount = ncount;
        Lv.npad = npad;
        Lv.type_bits = type_bits;
        Lv.tpa = tpa;
        Lv.built = built ? 1 : 0;
        Lv.wcap = tpa > 0 ? cap / tpa : 0;
        return Lv;
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
};

namespace {

void bdestroy(parm_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    bfree(b->geom);
    bfree(b->pos);
    bfree(b->pos_alt);
    bfree(b->xref);
    for (int k = 0; k < 3; k++) {
        bfree(b->v[k]);
        bfree(b->a[k]);
        bfree(b->f[k]);
    }
    bfree(b->m);
    bfree(b->diam);
    bfree(b->disp);
    bfree(b->stage);
    bfree(b->type);
    bfree(b->sys);
    bfree(b->nbr);
    bfree(b->ncount);
    bfree(b->keys);
    bfree(b->cellidx);
    bfree(b->xs);
    bfree(b->dsort);
    bfree(b->cstart);
    bfree(b->flags);
    bfree(b->work);
    bfree(b->wslice);
    bfree(b->which);
    bfree(b->gflag);
    bfree(b->nent);
    bfree(b->tab);
    bfree(b->eatom);
    bfree(b->sysout);
    for (auto &p : b->prof_force) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    for (auto &p : b->prof_rebuild) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    if (b->hgflag) (void)hipHostFree(b->hgflag);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

__global__ void kb_sys_index(int nsys, const BatchGeom *__restrict__ geom,
                             int32_t *__restrict__ sys) {
    const int s = blockIdx.x;
    if (s >= nsys) return;
    for (int k = threadIdx.x; k < geom[s].n; k += blockDim.x) sys[geom[s].off + k] = s;
}

__global__ void kb_upload(int ntot, int ndim, int what, BatchArrays A,
                          const double *__restrict__ src) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ntot) return;
    if (what == 0) {
        if (ndim == 3)
            static_cast<double4 *>(A.pos_out)[i] =
                double4{src[3 * (size_t)i], src[3 * (size_t)i + 1], src[3 * (size_t)i + 2], 0.0};
        else
            static_cast<double2 *>(A.pos_out)[i] =
                double2{src[2 * (size_t)i], src[2 * (size_t)i + 1]};
        return;
    }
    double *const *dst = what == 1 ? A.v : (what == 2 ? A.a : A.f);
    for (int k = 0; k < ndim; k++) dst[k][i] = src[(size_t)ndim * i + k];
}

__global__ void kb_download(int ntot, int ndim, int what, BatchArrays A,
                            double *__restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ntot) return;
    if (what == 0) {
        if (ndim == 3) {
            const double4 p = static_cast<const double4 *>(A.pos_in)[i];
            dst[3 * (size_t)i] = p.x;
            dst[3 * (size_t)i + 1] = p.y;
            dst[3 * (size_t)i + 2] = p.z;
        } else {
            const double2 p = static_cast<const double2 *>(A.pos_in)[i];
            dst[2 * (size_t)i] = p.x;
            dst[2 * (size_t)i + 1] = p.y;
        }
        return;
    }
    double *const *src = what == 1 ? A.v : (what == 2 ? A.a : A.f);
    for (int k = 0; k < ndim; k++) dst[(size_t)ndim * i + k] = src[k][i];
}

int nb256(long n) { return (int)((n + 255) / 256); }

int bupload(parm_batch *b, int what, const double *host) {
    const size_t bytes = sizeof(double) * (size_t)b->ntot * b->ndim;
    if (bytes == 0) return PARM_OK;
    PB_HIP(hipMemcpyAsync(b->stage, host, bytes, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(kb_upload, dim3(nb256(b->ntot)), dim3(256), 0, b->stream, b->ntot,
                       b->ndim, what, b->arrays(), b->stage);
    PB_HIP(hipGetLastError());
    PB_HIP(hipStreamSynchronize(b->stream));
    return PARM_OK;
}

int bdownload(parm_batch *b, int what, double *host) {
    const size_t bytes = sizeof(double) * (size_t)b->ntot * b->ndim;
    if (bytes == 0) return PARM_OK;
    hipLaunchKernelGGL(kb_download, dim3(nb256(b->ntot)), dim3(256), 0, b->stream, b->ntot,
                       b->ndim, what, b->arrays(), b->stage);
    PB_HIP(hipGetLastError());
    PB_HIP(hipMemcpyAsync(host, b->stage, bytes, hipMemcpyDeviceToHost, b->stream));
    PB_HIP(hipStreamSynchronize(b->stream));
    return PARM_OK;
}

int read_gflag(parm_batch *b, uint32_t *fl) {
    PB_HIP(hipMemcpyAsync(b->hgflag, b->gflag, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          b->stream));
    PB_HIP(hipStreamSynchronize(b->stream));
    *fl = *b->hgflag;
    return PARM_OK;
}

// Rebuild every system (force) or those whose skin flag is set.
// `alt`: positions one step ahead of `pos`; for a rebuilt system their
// displacement from the new reference positions is recomputed.
int brebuild(parm_batch *b, bool force, bool count, const void *pos,
             const void *alt = nullptr) {
    PB_HIP(launch_batch_rebuild(b->ndim, b->nsys, b->arrays(), pos, alt, force ? 1 : 0,
                                count ? 1 : 0, b->skin, b->type_bits, b->tpa, b->cap,
                                b->npad, b->nbr, b->ncount, b->keys, b->cellidx, b->cstart,
                                b->xs, b->dsort, b->flags, b->which, b->nent, b->gflag,
                                b->work, b->wslice, b->sort_grid, b->build_grid, b->stream));
    return PARM_OK;
}

int ensure_cap(parm_batch *b, int cap) {
    if (cap <= b->cap && b->nbr) return PARM_OK;
    bfree(b->nbr);
    b->cap = (cap + 7) & ~7;
    PB_CHECK(balloc(&b->nbr, (size_t)b->cap * b->npad));
    return PARM_OK;
}

// First build of every system (NeighborList::update_list on a new list),
// growing the list capacity until nothing overflows.
int first_build(parm_batch *b) {
    for (int attempt = 0; attempt < 6; attempt++) {
        PB_HIP(hipMemsetAsync(b->gflag, 0, sizeof(uint32_t), b->stream));
        PB_CHECK(brebuild(b, true, true, b->pos));
        uint32_t fl = 0;
        PB_CHECK(read_gflag(b, &fl));
        if (!(fl & kFlagOverflow)) {
            b->built = true;
            return PARM_OK;
        }
        // undo the counted rebuild, grow, redo
        std::vector<uint32_t> w(b->nsys);
        PB_HIP(hipMemcpy(w.data(), b->which, sizeof(uint32_t) * b->nsys,
                         hipMemcpyDeviceToHost));
        for (auto &x : w) x = x > 0 ? x - 1 : 0;
        PB_HIP(hipMemcpy(b->which, w.data(), sizeof(uint32_t) * b->nsys,
                         hipMemcpyHostToDevice));
        PB_CHECK(ensure_cap(b, b->cap * 2));
    }
    return bfail(PARM_E_DEVICE, "batch list build did not fit");
}

// Lanes-per-atom autotuner (as for one system): lay the lists out per
// candidate from the positions of the last rebuild (same pairs, `which`
// unchanged) and time force-only launches.
int bautotune(parm_batch *b) {
    if (b->tuned || b->tpa_fixed || !b->built || b->ntot == 0) return PARM_OK;
    hipEvent_t e0, e1;
    PB_HIP(hipEventCreate(&e0));
    PB_HIP(hipEventCreate(&e1));
    StepCoef c{};
    int best = b->tpa;
    float best_ms = 1e30f;
    for (int tpa : {1, 2, 4, 8}) {
        b->tpa = tpa;
        PB_CHECK(brebuild(b, true, false, b->xref));
        PB_HIP(launch_batch_force(b->ndim, b->prec, kStageForceOnly, false, b->arrays(),
                                  b->listview(), b->pairview(), c, b->eatom, b->stream));
        float tmin = 1e30f;
        for (int r = 0; r < 3; r++) {
            PB_HIP(hipEventRecord(e0, b->stream));
            PB_HIP(launch_batch_force(b->ndim, b->prec, kStageForceOnly, false, b->arrays(),
                                      b->listview(), b->pairview(), c, b->eatom, b->stream));
            PB_HIP(hipEventRecord(e1, b->stream));
            PB_HIP(hipEventSynchronize(e1));
            float ms = 0;
            PB_HIP(hipEventElapsedTime(&ms, e0, e1));
            tmin = std::min(tmin, ms);
        }
        if (tmin < best_ms * 0.97f) {
            best_ms = tmin;
            best = tpa;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tpa = best;
    PB_CHECK(brebuild(b, true, false, b->xref));
    b->tuned = true;
    return PARM_OK;
}

// One fused step.  The reference's timestep is drift, forces, kick, then
// update_trackers at the new positions.  Shifted by half a step: forces at
// x_q with the closing kick of step q-1 and the drift of step q into the other
// buffer (with |x_{q+1} - xref|), then update_trackers of step q-1: the
// rebuild at x_q of the systems whose skin test fired at x_q (recomputing
// |x_{q+1} - xref| for them), then the skin test at x_{q+1}.
int prof_begin(parm_batch *b, hipEvent_t *e) {
    e[0] = e[1] = nullptr;
    if (!b->prof) return PARM_OK;
    PB_HIP(hipEventCreate(&e[0]));
    PB_HIP(hipEventCreate(&e[1]));
    PB_HIP(hipEventRecord(e[0], b->stream));
    return PARM_OK;
}

int prof_end(parm_batch *b, hipEvent_t *e,
             std::vector<std::pair<hipEvent_t, hipEvent_t>> &dst) {
    if (!e[0]) return PARM_OK;
    PB_HIP(hipEventRecord(e[1], b->stream));
    dst.emplace_back(e[0], e[1]);
    return PARM_OK;
}

int fused_step(parm_batch *b, bool damped, const StepCoef &c, void *cur, void *nxt) {
    BatchArrays A = b->arrays();
    A.pos_in = cur;
    A.pos_out = nxt;
    hipEvent_t ev[2];
    PB_CHECK(prof_begin(b, ev));
    PB_HIP(launch_batch_force(b->ndim, b->prec,
                              damped ? kStageDampedKickDrift : kStageVerletKickDrift, false, A,
                              b->listview(), b->pairview(), c, b->eatom, b->stream));
    PB_CHECK(prof_end(b, ev, b->prof_force));
    PB_CHECK(prof_begin(b, ev));
    PB_CHECK(brebuild(b, false, true, cur, nxt));
    PB_CHECK(prof_end(b, ev, b->prof_rebuild));
    PB_HIP(launch_batch_skin(b->nsys, b->arrays(), b->skin, b->flags, b->gflag, b->stream));
    return PARM_OK;
}

int run_batch(parm_batch *b, bool damped, const StepCoef &c, int64_t nsteps, bool graph) {
    if (b->pair_kind == kPairNone) return bfail(PARM_E_STATE, "no pair interaction set");
    if (!b->built) return bfail(PARM_E_STATE, "call parm_batch_initialize first");
    if (nsteps <= 0) return PARM_OK;
    if (nsteps >= 16) PB_CHECK(bautotune(b));
    hipStream_t st = b->stream;
    PB_HIP(hipMemsetAsync(b->gflag, 0, sizeof(uint32_t), st));
    // first half of step 0 and its skin test
    PB_HIP(launch_batch_drift(b->ndim, damped, b->arrays(), c, st));
    PB_HIP(launch_batch_skin(b->nsys, b->arrays(), b->skin, b->flags, b->gflag, st));
    int64_t q = 0;
    const int64_t pairs = (nsteps - 1) / 2;
    if (graph && pairs > 0 && !b->prof) {
        const bool same = b->gexec && b->g_damped == damped && b->g_pos == b->pos &&
                          b->g_tpa == b->tpa && b->g_nbr == b->nbr && std::memcmp(&b->g_coef, &c, sizeof c) == 0;
        if (!same) {
            if (b->gexec) {
                (void)hipGraphExecDestroy(b->gexec);
                b->gexec = nullptr;
            }
            hipGraph_t g = nullptr;
            PB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc1 = fused_step(b, damped, c, b->pos, b->pos_alt);
            int rc2 = rc1 == PARM_OK ? fused_step(b, damped, c, b->pos_alt, b->pos) : rc1;
            hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc2 != PARM_OK) return rc2;
            if (ec != hipSuccess)
                return bfail(PARM_E_DEVICE, std::string("graph capture failed: ") +
                                                hipGetErrorString(ec));
            hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess)
                return bfail(PARM_E_DEVICE, std::string("graph instantiation failed: ") +
                                                hipGetErrorString(ei));
            b->g_damped = damped;
            b->g_pos = b->pos;
            b->g_tpa = b->tpa;
            b->g_nbr = b->nbr;
            b->g_coef = c;
        }
        for (int64_t r = 0; r < pairs; r++) PB_HIP(hipGraphLaunch(b->gexec, st));
        q = 2 * pairs;
    }
    for (; q < nsteps - 1; q++) {
        PB_CHECK(fused_step(b, damped, c, b->pos, b->pos_alt));
        std::swap(b->pos, b->pos_alt);
    }
    // last step: forces + closing kick, then update_trackers at x_n
    PB_HIP(launch_batch_force(b->ndim, b->prec, damped ? kStageDampedKick : kStageVerletKick,
                              false, b->arrays(), b->listview(), b->pairview(), c, b->eatom,
                              st));
    PB_CHECK(brebuild(b, false, true, b->pos));
    uint32_t fl = 0;
    PB_CHECK(read_gflag(b, &fl));
    if (fl & kFlagNonFinite) return bfail(PARM_E_NONFINITE, "a position became non-finite");
    if (fl & kFlagOverflow)
        return bfail(PARM_E_STATE,
                     "neighbour capacity exceeded during the run; results are invalid");
    return PARM_OK;
}

}  // namespace

extern "C" {

int parm_batch_create(int ndim, int precision, int64_t nsys, const int64_t *natoms,
                      const double *L, double skin, parm_batch **out) {
    if (!out) return bfail(PARM_E_INVALID, "out is null");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return bfail(PARM_E_INVALID, "ndim must be 2 or 3");
    if (precision != PARM_FP64 && precision != PARM_MIXED)
        return bfail(PARM_E_INVALID, "precision must be PARM_FP64 or PARM_MIXED");
    if (nsys <= 0 || !natoms || !L) return bfail(PARM_E_INVALID, "empty batch");
    if (!(skin >= 0) || !std::isfinite(skin)) return bfail(PARM_E_INVALID, "bad skin");
    int64_t tot = 0;
    for (int64_t s = 0; s < nsys; s++) {
        if (natoms[s] < 1) return bfail(PARM_E_INVALID, "every system needs atoms");
        tot += natoms[s];
        for (int k = 0; k < ndim; k++)
            if (!(L[s * ndim + k] > 0) || !std::isfinite(L[s * ndim + k]))
                return bfail(PARM_E_INVALID, "box edges must be finite and > 0");
    }
    if (tot >= (int64_t(1) << 28)) return bfail(PARM_E_INVALID, "too many atoms");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return bfail(PARM_E_DEVICE, "no HIP device available (the device path has no CPU fallback)");
    }
    parm_batch *b = new parm_batch();
    auto bail = [&](int r) {
        bdestroy(b);
        return r;
    };
    b->ndim = ndim;
    b->prec = precision;
    b->nsys = (int)nsys;
    b->ntot = (int)tot;
    b->npad = (int)((tot + 63) / 64 * 64);
    b->pos_bytes = ndim == 3 ? sizeof(double4) : sizeof(double2);
    b->skin = skin;
    b->off.resize(nsys + 1);
    b->off[0] = 0;
    for (int64_t s = 0; s < nsys; s++) b->off[s + 1] = b->off[s] + (int)natoms[s];
    b->L.assign(L, L + nsys * ndim);
    if (const char *e = std::getenv("PARM_TPA")) {
        int t = std::atoi(e);
        if (t == 1 || t == 2 || t == 4 || t == 8) {
            b->tpa = t;
            b->tpa_fixed = true;
        }
    }
    if (hipGetDevice(&b->device) != hipSuccess) return bail(bfail(PARM_E_DEVICE, "hipGetDevice"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return bail(bfail(PARM_E_DEVICE, "this library is built for gfx950 only"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(bfail(PARM_E_DEVICE, "stream creation failed"));
    // rebuild launches: the sort one 1024-thread workgroup per CU, the list
    // build eight 256-thread workgroups per CU (both loop over their work)
    b->sort_grid = std::max(1, prop.multiProcessorCount);
    b->build_grid = 8 * std::max(1, prop.multiProcessorCount);
    const size_t n = (size_t)tot;
    int rc;
    if ((rc = balloc((char **)&b->pos, n * b->pos_bytes)) ||
        (rc = balloc((char **)&b->pos_alt, n * b->pos_bytes)) ||
        (rc = balloc((char **)&b->xref, n * b->pos_bytes)))
        return bail(rc);
    for (int k = 0; k < 3; k++)
        if ((rc = balloc(&b->v[k], n)) || (rc = balloc(&b->a[k], n)) ||
            (rc = balloc(&b->f[k], n)))
            return bail(rc);
    if ((rc = balloc(&b->m, n)) || (rc = balloc(&b->diam, n)) || (rc = balloc(&b->disp, n)) ||
        (rc = balloc(&b->stage, n * ndim)) || (rc = balloc(&b->type, n)) ||
        (rc = balloc(&b->sys, n)) || (rc = balloc(&b->ncount, (size_t)b->npad)) ||
        (rc = balloc(&b->keys, n)) || (rc = balloc(&b->cellidx, n)) ||
        (rc = balloc((char **)&b->xs, n * b->pos_bytes)) || (rc = balloc(&b->dsort, n)) ||
        (rc = balloc(&b->flags, (size_t)nsys)) || (rc = balloc(&b->which, (size_t)nsys)) ||
        (rc = balloc(&b->work, (size_t)nsys + 1)) || (rc = balloc(&b->wslice, (size_t)nsys + 1)) ||
        (rc = balloc(&b->gflag, 1)) || (rc = balloc(&b->nent, (size_t)nsys)) ||
        (rc = balloc(&b->eatom, 4 * n)) || (rc = balloc(&b->sysout, 4 * (size_t)nsys)) ||
        (rc = balloc(&b->geom, (size_t)nsys)))
        return bail(rc);
    if (hipHostMalloc((void **)&b->hgflag, sizeof(uint32_t), 0) != hipSuccess)
        return bail(bfail(PARM_E_NOMEM, "pinned allocation failed"));
    hipStream_t st = b->stream;
    for (int k = 0; k < 3; k++) {
        (void)hipMemsetAsync(b->v[k], 0, n * sizeof(double), st);
        (void)hipMemsetAsync(b->a[k], 0, n * sizeof(double), st);
        (void)hipMemsetAsync(b->f[k], 0, n * sizeof(double), st);
    }
    (void)hipMemsetAsync(b->pos, 0, n * b->pos_bytes, st);
    (void)hipMemsetAsync(b->pos_alt, 0, n * b->pos_bytes, st);
    (void)hipMemsetAsync(b->xref, 0, n * b->pos_bytes, st);
    (void)hipMemsetAsync(b->disp, 0, n * sizeof(double), st);
    (void)hipMemsetAsync(b->type, 0, n * sizeof(int32_t), st);
    (void)hipMemsetAsync(b->flags, 0, nsys * sizeof(uint32_t), st);
    (void)hipMemsetAsync(b->which, 0, nsys * sizeof(uint32_t), st);
    (void)hipMemsetAsync(b->nent, 0, nsys * sizeof(unsigned long long), st);
    (void)hipMemsetAsync(b->gflag, 0, sizeof(uint32_t), st);
    std::vector<double> ones(n, 1.0);
    if (hipMemcpyAsync(b->m, ones.data(), n * sizeof(double), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return bail(bfail(PARM_E_DEVICE, "initialisation failed"));
    *out = b;
    return PARM_OK;
}

int parm_batch_destroy(parm_batch *b) {
    bdestroy(b);
    return PARM_OK;
}

int parm_batch_info(const parm_batch *b, int *ndim, int *nsys, int64_t *natoms_total) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    if (ndim) *ndim = b->ndim;
    if (nsys) *nsys = b->nsys;
    if (natoms_total) *natoms_total = b->ntot;
    return PARM_OK;
}

int parm_batch_set_masses(parm_batch *b, const double *masses) {
    if (!b || !masses) return bfail(PARM_E_INVALID, "null argument");
    PB_HIP(hipSetDevice(b->device));
    PB_HIP(hipMemcpy(b->m, masses, sizeof(double) * b->ntot, hipMemcpyHostToDevice));
    return PARM_OK;
}

int parm_batch_set_state(parm_batch *b, const double *x, const double *v, const double *a) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    PB_HIP(hipSetDevice(b->device));
    if (x) PB_CHECK(bupload(b, 0, x));
    if (v) PB_CHECK(bupload(b, 1, v));
    if (a) PB_CHECK(bupload(b, 2, a));
    return PARM_OK;
}

int parm_batch_get_state(parm_batch *b, double *x, double *v, double *a, double *f) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    PB_HIP(hipSetDevice(b->device));
    if (x) PB_CHECK(bdownload(b, 0, x));
    if (v) PB_CHECK(bdownload(b, 1, v));
    if (a) PB_CHECK(bdownload(b, 2, a));
    if (f) PB_CHECK(bdownload(b, 3, f));
    return PARM_OK;
}

static int batch_pair(parm_batch *b, int kind, int ntypes, const double *eps,
                      const double *sigma, const double *third, const int32_t *atype) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    if (ntypes < 1 || ntypes > 64) return bfail(PARM_E_UNSUPPORTED, "ntypes must be in [1, 64]");
    if (!eps || !sigma || !third || !atype) return bfail(PARM_E_INVALID, "null parameter array");
    for (int i = 0; i < b->ntot; i++)
        if (atype[i] < 0 || atype[i] >= ntypes)
            return bfail(PARM_E_INVALID, "atom type out of range");
    const std::string asym = check_rows_symmetric(kind, ntypes, eps, sigma);
    if (!asym.empty())
        return bfail(PARM_E_UNSUPPORTED,
                     asym + ": the device evaluates each pair from both of its atoms");
    std::vector<PairParam> tab;
    bool harm = true;
    make_pair_table(kind, ntypes, eps, sigma, third, tab, harm);
    // per-atom list diameters (A::max_size()) and per-system cell grids
    std::vector<double> dia(b->ntot);
    double dmax = 0;
    for (int i = 0; i < b->ntot; i++) {
        dia[i] = type_diameter(kind, sigma, third, atype[i]);
        dmax = std::max(dmax, dia[i]);
    }
    b->geom_h.resize(b->nsys);
    int cells = 0;
    for (int s = 0; s < b->nsys; s++) {
        BatchGeom G{};
        double smax = 0;
        for (int i = b->off[s]; i < b->off[s + 1]; i++) smax = std::max(smax, dia[i]);
        const double rmax = (smax + b->skin) * (1.0 + 1e-9);
        for (int k = 0; k < 3; k++) {
            const double Lk = k < b->ndim ? b->L[(size_t)s * b->ndim + k] : 1.0;
            G.box.L[k] = Lk;
            G.box.invL[k] = 1.0 / Lk;
            G.box.halfL[k] = Lk / 2;
            G.grid.nc[k] = 1;
            G.grid.invL[k] = k < b->ndim ? 1.0 / Lk : 0.0;
            if (k < b->ndim) {
                const double q = rmax > 0 ? std::floor(Lk / rmax) : 1.0;
                G.grid.nc[k] = q >= 3.0 ? (int)std::min(q, 4096.0) : 1;
            }
        }
        for (;;) {
            long tot = 1;
            for (int k = 0; k < b->ndim; k++) tot *= G.grid.nc[k];
            if (tot <= kBatchMaxCellsPerSystem) {
                G.grid.ncell = (int)tot;
                break;
            }
            int kk = 0;
            for (int k = 1; k < b->ndim; k++)
                if (G.grid.nc[k] > G.grid.nc[kk]) kk = k;
            G.grid.nc[kk] = G.grid.nc[kk] / 2 >= 3 ? G.grid.nc[kk] / 2 : 1;
        }
        G.off = b->off[s];
        G.n = b->off[s + 1] - b->off[s];
        G.cell_off = cells;
        cells += G.grid.ncell + 1;
        b->geom_h[s] = G;
    }
    PB_HIP(hipSetDevice(b->device));
    bfree(b->cstart);
    PB_CHECK(balloc(&b->cstart, (size_t)cells));
    b->total_cells = cells;
    bfree(b->tab);
    PB_CHECK(balloc(&b->tab, tab.size()));
    PB_HIP(hipMemcpy(b->tab, tab.data(), sizeof(PairParam) * tab.size(),
                     hipMemcpyHostToDevice));
    b->tab0 = tab[0];
    PB_HIP(hipMemcpy(b->geom, b->geom_h.data(), sizeof(BatchGeom) * b->nsys,
                     hipMemcpyHostToDevice));
    PB_HIP(hipMemcpy(b->type, atype, sizeof(int32_t) * b->ntot, hipMemcpyHostToDevice));
    PB_HIP(hipMemcpy(b->diam, dia.data(), sizeof(double) * b->ntot, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(kb_sys_index, dim3(b->nsys), dim3(256), 0, b->stream, b->nsys, b->geom,
                       b->sys);
    PB_HIP(hipGetLastError());
    PB_HIP(hipStreamSynchronize(b->stream));
    int bits = 0;
    while ((1 << bits) < ntypes) bits++;
    if (bits > 0 && (int64_t)b->ntot >= (int64_t(1) << (32 - bits)))
        return bfail(PARM_E_UNSUPPORTED, "too many atoms for this many types");
    b->type_bits = bits;
    b->ntypes = ntypes;
    b->single = ntypes == 1 ? 1 : 0;
    b->harmonic = (kind == kPairRepulsion && harm) ? 1 : 0;
    b->pair_kind = kind;
    b->built = false;
    // initial list capacity from the densest system
    double dens = 0;
    for (int s = 0; s < b->nsys; s++) {
        double vol = 1;
        for (int k = 0; k < b->ndim; k++) vol *= b->L[(size_t)s * b->ndim + k];
        dens = std::max(dens, (b->off[s + 1] - b->off[s]) / vol);
    }
    const double r = dmax + b->skin;
    const double ball = b->ndim == 3 ? 4.0 / 3.0 * kPi * r * r * r : kPi * r * r;
    PB_CHECK(ensure_cap(b, std::max(16, (int)(1.6 * dens * ball + 24))));
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    return PARM_OK;
}

int parm_batch_pair_repulsion(parm_batch *b, int ntypes, const double *eps,
                              const double *sigma, const double *exponent,
                              const int32_t *atom_type) {
    return batch_pair(b, kPairRepulsion, ntypes, eps, sigma, exponent, atom_type);
}

int parm_batch_pair_lj_attract_repulse(parm_batch *b, int ntypes, const double *eps_rows,
                                       const double *sigma, const double *sigcut,
                                       const int32_t *atom_type) {
    return batch_pair(b, kPairLJ, ntypes, eps_rows, sigma, sigcut, atom_type);
}

int parm_batch_initialize(parm_batch *b) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    if (b->pair_kind == kPairNone) return bfail(PARM_E_STATE, "no pair interaction set");
    PB_HIP(hipSetDevice(b->device));
    // Collection::initialize: update_list(false) (a new list: rebuild),
    // set_forces(true), update_list(false)
    PB_CHECK(first_build(b));
    StepCoef c{};
    PB_HIP(launch_batch_force(b->ndim, b->prec, kStageForceA, false, b->arrays(),
                              b->listview(), b->pairview(), c, b->eatom, b->stream));
    PB_HIP(launch_batch_disp(b->ndim, b->arrays(), b->stream));
    PB_HIP(launch_batch_skin(b->nsys, b->arrays(), b->skin, b->flags, b->gflag, b->stream));
    PB_CHECK(brebuild(b, false, true, b->pos));
    PB_HIP(hipStreamSynchronize(b->stream));
    return PARM_OK;
}

int parm_batch_damped_steps(parm_batch *b, double dt, double damping, int64_t nsteps,
                            int use_graph) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    if (dt <= 0) return bfail(PARM_E_INVALID, "Collection::CollectionSol: dt >= 0");
    PB_HIP(hipSetDevice(b->device));
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
    return run_batch(b, true, c, nsteps, use_graph != 0);
}

int parm_batch_verlet_steps(parm_batch *b, double dt, int64_t nsteps, int use_graph) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    PB_HIP(hipSetDevice(b->device));
    StepCoef c{};
    c.k0 = dt;
    c.k1 = dt * dt / 2;
    c.k2 = dt / 2;
    return run_batch(b, false, c, nsteps, use_graph != 0);
}

int parm_batch_pair_sums(parm_batch *b, double *U, double *W, double *contacts,
                         double *overlaps) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    if (b->pair_kind == kPairNone) return bfail(PARM_E_STATE, "no pair interaction set");
    PB_HIP(hipSetDevice(b->device));
    StepCoef c{};
    PB_HIP(launch_batch_force(b->ndim, b->prec, kStageSums, true, b->arrays(), b->listview(),
                              b->pairview(), c, b->eatom, b->stream));
    PB_HIP(launch_batch_sys_sums(b->nsys, b->arrays(), b->eatom, b->sysout, b->stream));
    double *dst[4] = {U, W, contacts, overlaps};
    for (int q = 0; q < 4; q++)
        if (dst[q])
            PB_HIP(hipMemcpyAsync(dst[q], b->sysout + (size_t)q * b->nsys,
                                  sizeof(double) * b->nsys, hipMemcpyDeviceToHost, b->stream));
    PB_HIP(hipStreamSynchronize(b->stream));
    return PARM_OK;
}

int parm_batch_sums(parm_batch *b, double *U, double *contacts) {
    return parm_batch_pair_sums(b, U, nullptr, contacts, nullptr);
}

int parm_batch_autotune(parm_batch *b) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    if (b->pair_kind == kPairNone) return bfail(PARM_E_STATE, "no pair interaction set");
    if (!b->built) return bfail(PARM_E_STATE, "call parm_batch_initialize first");
    PB_HIP(hipSetDevice(b->device));
    PB_CHECK(bautotune(b));
    PB_HIP(hipStreamSynchronize(b->stream));
    return PARM_OK;
}

int parm_batch_profile_enable(parm_batch *b, int enable) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    b->prof = enable != 0;
    return PARM_OK;
}

int parm_batch_profile_read(parm_batch *b, int reset, double *force_ms,
                            int64_t *force_launches, double *rebuild_ms,
                            int64_t *rebuild_launches) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    PB_HIP(hipSetDevice(b->device));
    PB_HIP(hipStreamSynchronize(b->stream));
    auto total = [](const std::vector<std::pair<hipEvent_t, hipEvent_t>> &v, double *ms) {
        double t = 0.0;
        for (const auto &p : v) {
            float e = 0.0f;
            if (hipEventElapsedTime(&e, p.first, p.second) == hipSuccess) t += e;
        }
        if (ms) *ms = t;
    };
    total(b->prof_force, force_ms);
    total(b->prof_rebuild, rebuild_ms);
    if (force_launches) *force_launches = (int64_t)b->prof_force.size();
    if (rebuild_launches) *rebuild_launches = (int64_t)b->prof_rebuild.size();
    if (reset) {
        for (auto *v : {&b->prof_force, &b->prof_rebuild}) {
            for (auto &p : *v) {
                (void)hipEventDestroy(p.first);
                (void)hipEventDestroy(p.second);
            }
            v->clear();
        }
    }
    return PARM_OK;
}

int parm_batch_lists(parm_batch *b, uint32_t *which, int64_t *npairs) {
    if (!b) return bfail(PARM_E_INVALID, "null batch");
    PB_HIP(hipSetDevice(b->device));
    if (which)
        PB_HIP(hipMemcpyAsync(which, b->which, sizeof(uint32_t) * b->nsys,
                              hipMemcpyDeviceToHost, b->stream));
    std::vector<unsigned long long> np(b->nsys);
    PB_HIP(hipMemcpyAsync(np.data(), b->nent, sizeof(unsigned long long) * b->nsys,
                          hipMemcpyDeviceToHost, b->stream));
    PB_HIP(hipStreamSynchronize(b->stream));
    if (npairs)
        for (int s = 0; s < b->nsys; s++) npairs[s] = (int64_t)(np[s] / 2);
    return PARM_OK;
}

int parm_batch_pairs(parm_batch *b, int64_t sysid, int64_t cap, int32_t *first,
                     int32_t *last, int64_t *npairs) {
    if (!b || sysid < 0 || sysid >= b->nsys) return bfail(PARM_E_INVALID, "bad system");
    if (cap > 0 && (!first || !last)) return bfail(PARM_E_INVALID, "null output");
    PB_HIP(hipSetDevice(b->device));
    const int o = b->off[sysid], n = b->off[sysid + 1] - o;
    std::vector<int32_t> cnt(n);
    std::vector<uint32_t> nb((size_t)b->cap * b->npad);
    PB_HIP(hipMemcpyAsync(cnt.data(), b->ncount + o, sizeof(int32_t) * n,
                          hipMemcpyDeviceToHost, b->stream));
    PB_HIP(hipMemcpyAsync(nb.data(), b->nbr, sizeof(uint32_t) * nb.size(),
                          hipMemcpyDeviceToHost, b->stream));
    PB_HIP(hipStreamSynchronize(b->stream));
    int64_t w = 0;
    std::vector<int32_t> lower;
    for (int li = 0; li < n; li++) {
        const int i = o + li;
        lower.clear();
        for (int q = 0; q < cnt[li]; q++) {
            const size_t idx = ((size_t)(q / b->tpa) * b->npad + i) * b->tpa + q % b->tpa;
            const int j = (int)(nb[idx] >> b->type_bits) - o;
            if (j < li) lower.push_back(j);
        }
        std::sort(lower.begin(), lower.end());
        for (int32_t lj : lower) {
            if (w < cap) {
                first[w] = li;
                last[w] = lj;
            }
            w++;
        }
    }
    if (npairs) *npairs = w;
    return PARM_OK;
}

}  // extern "C"
