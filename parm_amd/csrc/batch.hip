// batch.hip -- many independent systems (packings) stepped together.
//
// Each system is what the reference builds with one AtomVec, OriginBox,
// NeighborList, NListed pair interaction and CollectionDamped /
// CollectionVerlet (src/collection.cpp:356-381, 442-469; src/trackers.cpp:
// 19-85).  The systems' atoms are concatenated; every per-step kernel covers
// all systems at once, and every decision (skin test, rebuild) is taken on
// the device per system, so a step needs no host round trip and a run of
// steps can be captured in a HIP graph.
//
// Per step:  k_batch_force    forces + kick (+ next drift, |x - xref| per atom)
//            k_batch_worklist the systems whose skin flag is set, as a list
//            k_batch_sort     a fixed grid taking the listed systems in turn:
//                             cell sort in LDS, sorted copies, xref = x
//            k_batch_build    a fixed grid over slices of the listed
//                             systems: the exact list build, a thread per atom
//            k_batch_skin     one wavefront per system: top-2 of |x - xref|
//                             -> the system's flag for the next rebuild
#include "common.hpp"
#include "kernels.hpp"
#include "pairs.hpp"

namespace parm {

// ---------------------------------------------------------------- forces
constexpr int kBatchAhead = 4;  // list entries per lane evaluated before the loop

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE, bool ENERGY,
          int TPA>
__global__ __launch_bounds__(kBlock) void k_batch_force(BatchArrays A, ListView Lv,
                                                        PairView P, StepCoef c,
                                                        double *__restrict__ eatom) {
    using PT = typename PosStore<D>::T;
    constexpr bool DRIFT =
        (STAGE == kStageVerletKickDrift || STAGE == kStageDampedKickDrift);
    constexpr bool DAMPED =
        (STAGE == kStageDampedKick || STAGE == kStageDampedKickDrift);
    // consecutive workgroups (consecutive atoms: one system's, mostly) on one
    // XCD, so each system's positions, which its atoms' neighbour gathers
    // read, are held in one L2 rather than in all eight
    const int gt = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int i = gt / TPA;
    const int sub = gt % TPA;
    const bool act = i < A.ntot;
    PairAcc acc;
    double fi[D];
    float f32[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        fi[k] = 0.0;
        f32[k] = 0.0f;
    }
    const PT *pos = static_cast<const PT *>(A.pos_in);
    Vec<D> xi;
    Box box{};
    if (act) {
        xi = PosStore<D>::load(pos[i]);
        box = A.geom[A.sys[i]].box;
    }
    // the integrator inputs of the atom (lane sub == 0 applies the stage),
    // loaded now to arrive during the pair loop
    double m_i = 1.0, v_i[D], x0_i[D];
#pragma unroll
    for (int k = 0; k < D; k++) v_i[k] = x0_i[k] = 0.0;
    if (act && sub == 0 && STAGE != kStageSums && STAGE != kStageForceOnly) {
        m_i = A.m[i];
        if (STAGE != kStageForceA) {
#pragma unroll
            for (int k = 0; k < D; k++) v_i[k] = A.v[k][i];
        }
        if (DRIFT) {
            const Vec<D> x0 = PosStore<D>::load(static_cast<const PT *>(A.xref)[i]);
#pragma unroll
            for (int k = 0; k < D; k++) x0_i[k] = x0.c[k];
        }
    }
    if (act && Lv.built) {
        const size_t stride = (size_t)Lv.npad * TPA;
        const uint32_t *nb = Lv.nbr + (size_t)i * TPA + sub;
        // the lane's first kBatchAhead entries: their list words are read
        // together with the entry count (any word within the list array's
        // wcap per lane is in bounds; those past the count are not used),
        // then their neighbours' positions are gathered together
        uint32_t w[kBatchAhead];
#pragma unroll
        for (int k = 0; k < kBatchAhead; k++)
            w[k] = k < Lv.wcap ? nb[(size_t)k * stride] : 0u;
        const int cnt = Lv.ncount[i];
        const int tb = Lv.type_bits;
        const uint32_t tmask = (1u << tb) - 1u;
        const PairParam *row = P.tab;
        PairParam p0;
        if (SINGLE)
            p0 = P.p0;
        else
            row = P.tab + (size_t)A.type[i] * P.ntypes;
        PT pw[kBatchAhead];
#pragma unroll
        for (int k = 0; k < kBatchAhead; k++)
            if (sub + k * TPA < cnt) pw[k] = pos[w[k] >> tb];
#pragma unroll
        for (int k = 0; k < kBatchAhead; k++) {
            if (sub + k * TPA < cnt) {
                const Vec<D> r = diff<D>(xi, PosStore<D>::load(pw[k]), box);
                visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : row[w[k] & tmask], r, fi,
                                                   f32, acc);
            }
        }
        // the rest: the next two entries loaded while the current two are
        // evaluated
        int q = sub + kBatchAhead * TPA;
        size_t off = (size_t)kBatchAhead * stride;
        uint32_t n0 = q < cnt ? nb[off] : 0u;
        uint32_t n1 = q + TPA < cnt ? nb[off + stride] : 0u;
        for (; q + TPA < cnt; q += 2 * TPA, off += 2 * stride) {
            const uint32_t e0 = n0;
            const uint32_t e1 = n1;
            if (q + 2 * TPA < cnt) n0 = nb[off + 2 * stride];
            if (q + 3 * TPA < cnt) n1 = nb[off + 3 * stride];
            const PT pj0 = pos[e0 >> tb];
            const PT pj1 = pos[e1 >> tb];
            const Vec<D> r0 = diff<D>(xi, PosStore<D>::load(pj0), box);
            visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : row[e0 & tmask], r0, fi, f32,
                                               acc);
            const Vec<D> r1 = diff<D>(xi, PosStore<D>::load(pj1), box);
            visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : row[e1 & tmask], r1, fi, f32,
                                               acc);
        }
        if (q < cnt) {
            const uint32_t e0 = n0;
            const Vec<D> r0 = diff<D>(xi, PosStore<D>::load(pos[e0 >> tb]), box);
            visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : row[e0 & tmask], r0, fi, f32,
                                               acc);
        }
    }
    if (PREC != kFP64) {
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] = (double)f32[k];
    }
    double e = acc.e, nc = acc.contacts, w = acc.w, no = acc.overlaps;
#pragma unroll
    for (int o = 1; o < TPA; o <<= 1) {
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] += __shfl_xor(fi[k], o);
        if (ENERGY) {
            e += __shfl_xor(e, o);
            nc += __shfl_xor(nc, o);
            w += __shfl_xor(w, o);
            no += __shfl_xor(no, o);
        }
    }
    if (!act || sub != 0) return;
    if (ENERGY) {
        // the full list visits each pair from both atoms: halves per atom
        double4 *ea = reinterpret_cast<double4 *>(eatom);
        ea[i] = double4{0.5 * e, nc, 0.5 * w, no};
    }
    if (STAGE == kStageSums) return;
    // fused kick + drift: f and a are stored by the run's last (plain kick)
    // launch only; nothing reads them in between
    if constexpr (!DRIFT) {
#pragma unroll
        for (int k = 0; k < D; k++) A.f[k][i] = fi[k];
    }
    if (STAGE == kStageForceOnly) return;
    const double m = m_i;
    const bool fixed = is_fixed(m);
    double a[D];
#pragma unroll
    for (int k = 0; k < D; k++) a[k] = fixed ? 0.0 : fi[k] / m;
    if constexpr (!DRIFT) {
#pragma unroll
        for (int k = 0; k < D; k++) A.a[k][i] = a[k];
    }
    if (STAGE == kStageForceA) return;
    double v[D];
#pragma unroll
    for (int k = 0; k < D; k++) v[k] = v_i[k];
    if (!fixed) {
#pragma unroll
        for (int k = 0; k < D; k++) v[k] = v[k] + a[k] * (DAMPED ? c.k4 : c.k2);
    }
    if (DRIFT) {
        Vec<D> xn = xi;
        drift_atom<D, DAMPED>(xn, v, a, fixed, c);
        static_cast<PT *>(A.pos_out)[i] = PosStore<D>::make(xn);
        Vec<D> d;
#pragma unroll
        for (int k = 0; k < D; k++) d.c[k] = xn.c[k] - x0_i[k];
        A.disp[i] = sqrt(sqnorm<D>(d));
    }
#pragma unroll
    for (int k = 0; k < D; k++) A.v[k][i] = v[k];
}

// First half of a step (drift + half kick) and |x - xref|.
template <int D, bool DAMPED>
__global__ __launch_bounds__(kBlock) void k_batch_drift(BatchArrays A, StepCoef c) {
    using PT = typename PosStore<D>::T;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= A.ntot) return;
    PT *pos = static_cast<PT *>(A.pos_out);
    Vec<D> x = PosStore<D>::load(pos[i]);
    double v[D], a[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        v[k] = A.v[k][i];
        a[k] = A.a[k][i];
    }
    const bool fixed = is_fixed(A.m[i]);
    drift_atom<D, DAMPED>(x, v, a, fixed, c);
    if (!fixed) pos[i] = PosStore<D>::make(x);
#pragma unroll
    for (int k = 0; k < D; k++) A.v[k][i] = v[k];
    const Vec<D> x0 = PosStore<D>::load(static_cast<const PT *>(A.xref)[i]);
    Vec<D> d;
#pragma unroll
    for (int k = 0; k < D; k++) d.c[k] = x.c[k] - x0.c[k];
    A.disp[i] = sqrt(sqnorm<D>(d));
}

// |x - xref| for every atom (update_list(false) outside a step).
template <int D>
__global__ __launch_bounds__(kBlock) void k_batch_disp(BatchArrays A) {
    using PT = typename PosStore<D>::T;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= A.ntot) return;
    const Vec<D> x = PosStore<D>::load(static_cast<const PT *>(A.pos_in)[i]);
    const Vec<D> x0 = PosStore<D>::load(static_cast<const PT *>(A.xref)[i]);
    Vec<D> d;
#pragma unroll
    for (int k = 0; k < D; k++) d.c[k] = x.c[k] - x0.c[k];
    A.disp[i] = sqrt(sqnorm<D>(d));
}

// ------------------------------------------------------------ skin test
// One wavefront per system: the reference's rule (src/trackers.cpp:23-53),
// rebuild iff the two largest displacements sum to >= skin, checked only
// once something moved.  The flag is overwritten, never accumulated.
__global__ __launch_bounds__(kBlock) void k_batch_skin(int nsys, BatchArrays A,
                                                       double skin,
                                                       uint32_t *__restrict__ flags,
                                                       uint32_t *__restrict__ gflag) {
    const int s = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nsys) return;
    const int off = A.geom[s].off, n = A.geom[s].n;
    Top2 t;
    bool bad = false;
    for (int k = lane; k < n; k += kWave) {
        const double d = A.disp[off + k];
        t.add(d);
        bad |= !isfinite(d);
    }
    t = wave_top2(t);
    const unsigned long long anybad = __ballot(bad);
    if (lane == 0) {
        flags[s] = (t.a > 0 && t.b + t.a >= skin) ? 1u : 0u;
        if (anybad) atomicOr(gflag, kFlagNonFinite);
    }
}

// -------------------------------------------------------------- rebuild
// The systems to rebuild at the end of a step are those whose skin flag is
// set (or all, when forced).  k_batch_worklist turns the flags into a list
// (system ids in ascending order, count first) and counts each listed
// system's slices of kBuildSlice atoms (exclusive prefix in wslice).  Then
// two launches on fixed grids, so the launch shapes do not depend on how
// many systems rebuild (the step stays graph-capturable) and no workgroup is
// started for a system that keeps its list:
//   k_batch_sort   one workgroup of kSortThreads per listed system: cell
//                  keys, a stable counting sort of the system's atoms by
//                  cell in LDS, the cell starts, sorted copies of the
//                  positions and diameters, xref = x;
//   k_batch_build  one workgroup of kBuildSlice threads per slice of a
//                  listed system, a thread per atom: the exact list build
//                  over the periodic 27-cell stencil (candidates in cell
//                  order, then sorted order: deterministic), four candidates
//                  in flight per thread.  A large system's build is spread
//                  over many workgroups.
// With `alt` (the positions one step ahead), |alt - xref| is recomputed for
// the next skin test.
constexpr int kBatchMaxCells = kBatchMaxCellsPerSystem;
constexpr int kSortThreads = 1024;
constexpr int kWorklistThreads = 1024;
constexpr int kBuildSlice = 256;

__global__ __launch_bounds__(kWorklistThreads) void k_batch_worklist(
    int nsys, const uint32_t *__restrict__ flags, int force,
    const BatchGeom *__restrict__ geom, int32_t *__restrict__ work,
    int32_t *__restrict__ wslice) {
    __shared__ int wsum[kWorklistThreads / kWave][2];
    const int t = threadIdx.x;
    const int per = (nsys + kWorklistThreads - 1) / kWorklistThreads;
    const int s0 = t * per;
    int loc = 0, lsl = 0;
    for (int q = 0; q < per; q++) {
        const int s = s0 + q;
        if (s < nsys && (force || flags[s] != 0u)) {
            loc++;
            lsl += (geom[s].n + kBuildSlice - 1) / kBuildSlice;
        }
    }
    const int lane = t % kWave, w = t / kWave;
    int incl = loc, isl = lsl;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o), z = __shfl_up(isl, o);
        if (lane >= o) {
            incl += y;
            isl += z;
        }
    }
    if (lane == kWave - 1) {
        wsum[w][0] = incl;
        wsum[w][1] = isl;
    }
    __syncthreads();
    int pos = incl - loc, psl = isl - lsl, tot = 0, totsl = 0;
    for (int q = 0; q < kWorklistThreads / kWave; q++) {
        if (q < w) {
            pos += wsum[q][0];
            psl += wsum[q][1];
        }
        tot += wsum[q][0];
        totsl += wsum[q][1];
    }
    for (int q = 0; q < per; q++) {
        const int s = s0 + q;
        if (s < nsys && (force || flags[s] != 0u)) {
            work[1 + pos] = s;
            wslice[pos] = psl;
            pos++;
            psl += (geom[s].n + kBuildSlice - 1) / kBuildSlice;
        }
    }
    if (t == 0) {
        work[0] = tot;
        wslice[tot] = totsl;
    }
}

template <int D>
__device__ __forceinline__ void sort_system(
    const int s, const BatchArrays &A, const typename PosStore<D>::T *__restrict__ pos,
    int count_rebuild, int32_t *__restrict__ keys, int32_t *__restrict__ cellidx,
    int32_t *__restrict__ cstart, typename PosStore<D>::T *__restrict__ xs,
    double *__restrict__ dsort, uint32_t *__restrict__ flags, uint32_t *__restrict__ which,
    unsigned long long *__restrict__ nent, int *hist) {
    using PT = typename PosStore<D>::T;
    constexpr int NT = kSortThreads;
    const BatchGeom G = A.geom[s];
    const int off = G.off, n = G.n, ncell = G.grid.ncell;
    const int t = threadIdx.x;
    for (int c = t; c < ncell; c += NT) hist[c] = 0;
    __syncthreads();
    // keys and counts
    for (int k = t; k < n; k += NT) {
        const Vec<D> x = PosStore<D>::load(pos[off + k]);
        int key = 0;
#pragma unroll
        for (int d = D - 1; d >= 0; d--) {
            double sv = x.c[d] * G.grid.invL[d];
            sv -= floor(sv);
            int cd = (int)(sv * (double)G.grid.nc[d]);
            cd = cd < 0 ? 0 : (cd >= G.grid.nc[d] ? G.grid.nc[d] - 1 : cd);
            key = key * G.grid.nc[d] + cd;
        }
        keys[off + k] = key;
        atomicAdd(&hist[key], 1);
    }
    __syncthreads();
    // exclusive scan of the counts: each thread sums a run of kScanPer
    // cells, a block scan of the run sums, then each thread writes its run
    {
        constexpr int kScanPer = kBatchMaxCells / NT;
        __shared__ int wsum[NT / kWave];
        const int c0 = t * kScanPer;
        int loc = 0;
        for (int q = 0; q < kScanPer; q++) loc += (c0 + q < ncell) ? hist[c0 + q] : 0;
        const int lane = t % kWave, w = t / kWave;
        int incl = loc;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == kWave - 1) wsum[w] = incl;
        __syncthreads();
        int run = incl - loc;
        for (int q = 0; q < w; q++) run += wsum[q];
        __syncthreads();
        for (int q = 0; q < kScanPer; q++) {
            const int c = c0 + q;
            if (c < ncell) {
                const int h = hist[c];
                cstart[G.cell_off + c] = off + run;
                hist[c] = run;
                run += h;
            }
        }
        if (t == 0) cstart[G.cell_off + ncell] = off + n;
    }
    __syncthreads();
    // stable scatter, one chunk of NT atoms at a time: within a wavefront the
    // rank among equal keys comes from ballots over the key bits; the
    // wavefronts of the chunk then take their turn, in order, on the cell
    // counters (the wavefront's reads of hist[key] precede, in its program
    // order, the update by its last lane of that key)
    {
        constexpr int kKeyBits = 12;
        static_assert(kBatchMaxCells <= (1 << kKeyBits), "cell keys fit the ballot match");
        const int lane = t % kWave, wv = t / kWave;
        const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        for (int base = 0; base < n; base += NT) {
            const int k = base + t;
            const bool valid = k < n;
            const int key = valid ? keys[off + k] : 0;
            unsigned long long eq = __ballot(valid);
#pragma unroll
            for (int b = 0; b < kKeyBits; b++) {
                const bool bit = (key >> b) & 1;
                const unsigned long long bb = __ballot(bit);
                eq &= bit ? bb : ~bb;
            }
            const int r = __popcll(eq & below);
            const bool lastk = valid && (eq >> lane) == 1ull;
            const int cntk = __popcll(eq);
            for (int w2 = 0; w2 < NT / kWave; w2++) {
                if (wv == w2) {
                    if (valid) cellidx[off + hist[key] + r] = off + k;
                    if (lastk) hist[key] += cntk;
                }
                __syncthreads();
            }
        }
    }
    // sorted copies of the positions and diameters, and xref = x (the
    // reference stores lastlocs inside the rebuild loop)
    for (int u = t; u < n; u += NT) {
        const int i = cellidx[off + u];
        xs[off + u] = pos[i];
        dsort[off + u] = A.diam[i];
    }
    for (int k = t; k < n; k += NT) static_cast<PT *>(A.xref_w)[off + k] = pos[off + k];
    if (t == 0) {
        nent[s] = 0ull;
        if (count_rebuild) which[s] += 1u;
        flags[s] = 0u;
    }
    // the next listed system reuses the LDS
    __syncthreads();
}

template <int D>
__global__ __launch_bounds__(kSortThreads) void k_batch_sort(
    BatchArrays A, const typename PosStore<D>::T *__restrict__ pos, int count_rebuild,
    int32_t *__restrict__ keys, int32_t *__restrict__ cellidx, int32_t *__restrict__ cstart,
    typename PosStore<D>::T *__restrict__ xs, double *__restrict__ dsort,
    uint32_t *__restrict__ flags, uint32_t *__restrict__ which,
    unsigned long long *__restrict__ nent, const int32_t *__restrict__ work) {
    __shared__ int hist[kBatchMaxCells];
    const int count = work[0];
    for (int q = blockIdx.x; q < count; q += gridDim.x)
        sort_system<D>(work[1 + q], A, pos, count_rebuild, keys, cellidx, cstart, xs, dsort,
                       flags, which, nent, hist);
}

template <int D>
__global__ __launch_bounds__(kBuildSlice) void k_batch_build(
    BatchArrays A, const typename PosStore<D>::T *__restrict__ alt, double skin,
    int type_bits, int tpa, int cap, int npad, uint32_t *__restrict__ nbr,
    int32_t *__restrict__ ncount, const int32_t *__restrict__ keys,
    const int32_t *__restrict__ cellidx, const int32_t *__restrict__ cstart,
    const typename PosStore<D>::T *__restrict__ xs, const double *__restrict__ dsort,
    unsigned long long *__restrict__ nent, uint32_t *__restrict__ gflag,
    const int32_t *__restrict__ work, const int32_t *__restrict__ wslice) {
    __shared__ unsigned long long red[kBuildSlice / kWave];
    const int count = work[0];
    const int total = wslice[count];
    const int t = threadIdx.x;
    for (int gs = blockIdx.x; gs < total; gs += gridDim.x) {
        // the listed system holding slice gs: the last q with wslice[q] <= gs
        int lo = 0, hi = count;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (wslice[mid] <= gs)
                lo = mid;
            else
                hi = mid;
        }
        const int s = work[1 + lo];
        const BatchGeom G = A.geom[s];
        const int off = G.off, n = G.n, nc0 = G.grid.nc[0], nc1 = G.grid.nc[1],
                  nc2 = G.grid.nc[2];
        const int u = (gs - wslice[lo]) * kBuildSlice + t;
        unsigned long long mycnt = 0;
        if (u < n) {
            const int lo0 = nc0 >= 3 ? -1 : 0, lo1 = (D >= 2 && nc1 >= 3) ? -1 : 0,
                      lo2 = (D == 3 && nc2 >= 3) ? -1 : 0;
            const int i = cellidx[off + u];
            const Vec<D> xi = PosStore<D>::load(xs[off + u]);
            const double di = dsort[off + u];
            int key = keys[i];
            int ci[3] = {0, 0, 0};
            ci[0] = key % nc0;
            key /= nc0;
            if (D >= 2) {
                ci[1] = key % nc1;
                key /= nc1;
            }
            if (D == 3) ci[2] = key % nc2;
            int cnt = 0;
            for (int oz = lo2; oz <= -lo2; oz++)
                for (int oy = lo1; oy <= -lo1; oy++)
                    for (int ox = lo0; ox <= -lo0; ox++) {
                        int q0 = ci[0] + ox, q1 = ci[1] + oy, q2 = ci[2] + oz;
                        q0 = q0 < 0 ? q0 + nc0 : (q0 >= nc0 ? q0 - nc0 : q0);
                        q1 = q1 < 0 ? q1 + nc1 : (q1 >= nc1 ? q1 - nc1 : q1);
                        q2 = q2 < 0 ? q2 + nc2 : (q2 >= nc2 ? q2 - nc2 : q2);
                        const int cc = (D == 3 ? (q2 * nc1 + q1) * nc0 : q1 * nc0) + q0;
                        const int s0 = cstart[G.cell_off + cc] - off;
                        const int s1 = cstart[G.cell_off + cc + 1] - off;
                        for (int v0 = s0; v0 < s1; v0 += 4) {
                            // four candidates' loads in flight together
                            Vec<D> xj[4];
                            double dj[4];
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                const int v = v0 + q < s1 ? v0 + q : s0;
                                xj[q] = PosStore<D>::load(xs[off + v]);
                                dj[q] = dsort[off + v];
                            }
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                const int v = v0 + q;
                                if (v >= s1 || v == u) continue;
                                // minimum image by rint: equal to the
                                // reference's remainder up to rounding far
                                // inside the 1e-6 band, where the reference's
                                // predicate decides
                                double d2 = 0.0;
#pragma unroll
                                for (int k = 0; k < D; k++) {
                                    double dk = xi.c[k] - xj[q].c[k];
                                    dk -= G.box.L[k] * rint(dk * G.box.invL[k]);
                                    d2 = fma(dk, dk, d2);
                                }
                                const double cut = (di + dj[q]) / 2 + skin;
                                const double c2 = cut * cut;
                                const bool pass =
                                    d2 < c2 * (1.0 - 1e-6) ||
                                    (d2 <= c2 * (1.0 + 1e-6) &&
                                     in_list<D>(xi, xj[q], di, dj[q], G.box, skin));
                                if (pass) {
                                    if (cnt < cap) {
                                        const int j = cellidx[off + v];
                                        nbr[untiled_entry(cnt, i, tpa, npad)] =
                                            ((uint32_t)j << type_bits) | (uint32_t)A.type[j];
                                    }
                                    cnt++;
                                }
                            }
                        }
                    }
            ncount[i] = cnt < cap ? cnt : cap;
            if (cnt > cap) atomicOr(gflag, kFlagOverflow);
            mycnt = (unsigned long long)cnt;
            if (alt) {
                const Vec<D> xa = PosStore<D>::load(alt[i]);
                Vec<D> dd;
#pragma unroll
                for (int d = 0; d < D; d++) dd.c[d] = xa.c[d] - xi.c[d];
                A.disp[i] = sqrt(sqnorm<D>(dd));
            }
        }
        unsigned long long sum = mycnt;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (t % kWave == 0) red[t / kWave] = sum;
        __syncthreads();
        if (t == 0) {
            unsigned long long tot = 0;
            for (int q = 0; q < kBuildSlice / kWave; q++) tot += red[q];
            atomicAdd(&nent[s], tot);
        }
        __syncthreads();
    }
}

// Per-system sums of the per-atom energies, virials and contact / overlap
// counts (NListed::energy, pressure, contacts, overlaps per system): one
// wavefront per system, compensated sums for the energy and the virial.
__global__ __launch_bounds__(kBlock) void k_batch_sys_sums(int nsys, BatchArrays A,
                                                           const double *__restrict__ eatom,
                                                           double *__restrict__ out) {
    const int s = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nsys) return;
    const int off = A.geom[s].off, n = A.geom[s].n;
    const double4 *ea = reinterpret_cast<const double4 *>(eatom);
    KSum e, w;
    double c = 0.0, o = 0.0;
    for (int k = lane; k < n; k += kWave) {
        const double4 q = ea[off + k];
        e.add(q.x);
        c += q.y;
        w.add(q.z);
        o += q.w;
    }
    CSum ce{e.s, e.c}, cw{w.s, w.c};
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        ce = csum_merge(ce, CSum{__shfl_xor(ce.s, d), __shfl_xor(ce.c, d)});
        cw = csum_merge(cw, CSum{__shfl_xor(cw.s, d), __shfl_xor(cw.c, d)});
        c += __shfl_xor(c, d);
        o += __shfl_xor(o, d);
    }
    if (lane == 0) {
        out[s] = ce.s + ce.c;
        out[nsys + s] = cw.s + cw.c;
        out[2 * (size_t)nsys + s] = c / 2;
        out[3 * (size_t)nsys + s] = o / 2;
    }
}

// ======================================================== host launchers
static inline int nblk1(long n) { return (int)((n + kBlock - 1) / kBlock); }

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE, int TPA>
static void bforce_tt(bool energy, const BatchArrays &A, const ListView &Lv,
                      const PairView &P, const StepCoef &c, double *eatom, hipStream_t s) {
    // energies are accumulated by the sums stage only
    (void)energy;
    dim3 g(nblk1((long)A.ntot * TPA)), b(kBlock);
    hipLaunchKernelGGL(
        (k_batch_force<D, KIND, PREC, SINGLE, HARM, STAGE, STAGE == kStageSums, TPA>), g, b,
        0, s, A, Lv, P, c, eatom);
}

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE>
static void bforce_t(bool energy, const BatchArrays &A, const ListView &Lv,
                     const PairView &P, const StepCoef &c, double *eatom, hipStream_t s) {
    switch (Lv.tpa) {
    case 1:
        bforce_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 1>(energy, A, Lv, P, c, eatom, s);
        break;
    case 2:
        bforce_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 2>(energy, A, Lv, P, c, eatom, s);
        break;
    case 8:
        bforce_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 8>(energy, A, Lv, P, c, eatom, s);
        break;
    default:
        bforce_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 4>(energy, A, Lv, P, c, eatom, s);
        break;
    }
}

template <int D, int KIND, int PREC, bool SINGLE, bool HARM>
static void bforce_s(int stage, bool energy, const BatchArrays &A, const ListView &Lv,
                     const PairView &P, const StepCoef &c, double *eatom, hipStream_t s) {
    switch (stage) {
#define PARM_BSTAGE(S)                                                          \
    case S:                                                                     \
        bforce_t<D, KIND, PREC, SINGLE, HARM, S>(energy, A, Lv, P, c, eatom, s); \
        break;
        PARM_BSTAGE(kStageForceOnly)
        PARM_BSTAGE(kStageForceA)
        PARM_BSTAGE(kStageVerletKick)
        PARM_BSTAGE(kStageVerletKickDrift)
        PARM_BSTAGE(kStageDampedKick)
        PARM_BSTAGE(kStageDampedKickDrift)
        PARM_BSTAGE(kStageSums)
#undef PARM_BSTAGE
    }
}

template <int D, int KIND, int PREC>
static void bforce_k(int stage, bool energy, const BatchArrays &A, const ListView &Lv,
                     const PairView &P, const StepCoef &c, double *eatom, hipStream_t s) {
    if constexpr (KIND == kPairRepulsion) {
        if (P.harmonic) {
            if (P.single)
                bforce_s<D, KIND, PREC, true, true>(stage, energy, A, Lv, P, c, eatom, s);
            else
                bforce_s<D, KIND, PREC, false, true>(stage, energy, A, Lv, P, c, eatom, s);
            return;
        }
    }
    {
        if (P.single)
            bforce_s<D, KIND, PREC, true, false>(stage, energy, A, Lv, P, c, eatom, s);
        else
            bforce_s<D, KIND, PREC, false, false>(stage, energy, A, Lv, P, c, eatom, s);
    }
}

template <int D>
static void bforce_d(int prec, int stage, bool energy, const BatchArrays &A,
                     const ListView &Lv, const PairView &P, const StepCoef &c,
                     double *eatom, hipStream_t s) {
    if (P.kind == kPairRepulsion) {
        if (prec == kFP64)
            bforce_k<D, kPairRepulsion, kFP64>(stage, energy, A, Lv, P, c, eatom, s);
        else
            bforce_k<D, kPairRepulsion, kMixed>(stage, energy, A, Lv, P, c, eatom, s);
    } else {
        if (prec == kFP64)
            bforce_k<D, kPairLJ, kFP64>(stage, energy, A, Lv, P, c, eatom, s);
        else
            bforce_k<D, kPairLJ, kMixed>(stage, energy, A, Lv, P, c, eatom, s);
    }
}

hipError_t launch_batch_force(int ndim, int precision, int stage, bool energy,
                              const BatchArrays &A, const ListView &Lv, const PairView &P,
                              const StepCoef &c, double *eatom, hipStream_t s) {
    if (A.ntot <= 0) return hipSuccess;
    if (ndim == 3)
        bforce_d<3>(precision, stage, energy, A, Lv, P, c, eatom, s);
    else
        bforce_d<2>(precision, stage, energy, A, Lv, P, c, eatom, s);
    return hipGetLastError();
}

hipError_t launch_batch_drift(int ndim, bool damped, const BatchArrays &A,
                              const StepCoef &c, hipStream_t s) {
    if (A.ntot <= 0) return hipSuccess;
    dim3 g(nblk1(A.ntot)), b(kBlock);
    if (ndim == 3) {
        if (damped)
            hipLaunchKernelGGL((k_batch_drift<3, true>), g, b, 0, s, A, c);
        else
            hipLaunchKernelGGL((k_batch_drift<3, false>), g, b, 0, s, A, c);
    } else {
        if (damped)
            hipLaunchKernelGGL((k_batch_drift<2, true>), g, b, 0, s, A, c);
        else
            hipLaunchKernelGGL((k_batch_drift<2, false>), g, b, 0, s, A, c);
    }
    return hipGetLastError();
}

hipError_t launch_batch_disp(int ndim, const BatchArrays &A, hipStream_t s) {
    if (A.ntot <= 0) return hipSuccess;
    dim3 g(nblk1(A.ntot)), b(kBlock);
    if (ndim == 3)
        hipLaunchKernelGGL(k_batch_disp<3>, g, b, 0, s, A);
    else
        hipLaunchKernelGGL(k_batch_disp<2>, g, b, 0, s, A);
    return hipGetLastError();
}

hipError_t launch_batch_skin(int nsys, const BatchArrays &A, double skin, uint32_t *flags,
                             uint32_t *gflag, hipStream_t s) {
    const int per = kBlock / kWave;
    hipLaunchKernelGGL(k_batch_skin, dim3((nsys + per - 1) / per), dim3(kBlock), 0, s, nsys,
                       A, skin, flags, gflag);
    return hipGetLastError();
}

hipError_t launch_batch_rebuild(int ndim, int nsys, const BatchArrays &A, const void *pos,
                                const void *alt, int force, int count_rebuild, double skin,
                                int type_bits, int tpa, int cap, int npad, uint32_t *nbr,
                                int32_t *ncount, int32_t *keys, int32_t *cellidx,
                                int32_t *cstart, void *xs, double *dsort, uint32_t *flags,
                                uint32_t *which, unsigned long long *nent, uint32_t *gflag,
                                int32_t *work, int32_t *wslice, int sort_grid, int build_grid,
                                hipStream_t s) {
    if (nsys <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_worklist, dim3(1), dim3(kWorklistThreads), 0, s, nsys, flags,
                       force, A.geom, work, wslice);
    const int gs = sort_grid < nsys ? sort_grid : nsys;
    if (ndim == 3) {
        hipLaunchKernelGGL(k_batch_sort<3>, dim3(gs), dim3(kSortThreads), 0, s, A,
                           (const double4 *)pos, count_rebuild, keys, cellidx, cstart,
                           (double4 *)xs, dsort, flags, which, nent, work);
        hipLaunchKernelGGL(k_batch_build<3>, dim3(build_grid), dim3(kBuildSlice), 0, s, A,
                           (const double4 *)alt, skin, type_bits, tpa, cap, npad, nbr, ncount,
                           keys, cellidx, cstart, (const double4 *)xs, dsort, nent, gflag,
                           work, wslice);
    } else {
        hipLaunchKernelGGL(k_batch_sort<2>, dim3(gs), dim3(kSortThreads), 0, s, A,
                           (const double2 *)pos, count_rebuild, keys, cellidx, cstart,
                           (double2 *)xs, dsort, flags, which, nent, work);
        hipLaunchKernelGGL(k_batch_build<2>, dim3(build_grid), dim3(kBuildSlice), 0, s, A,
                           (const double2 *)alt, skin, type_bits, tpa, cap, npad, nbr, ncount,
                           keys, cellidx, cstart, (const double2 *)xs, dsort, nent, gflag,
                           work, wslice);
    }
    return hipGetLastError();
}

hipError_t launch_batch_sys_sums(int nsys, const BatchArrays &A, const double *eatom,
                                 double *out, hipStream_t s) {
    const int per = kBlock / kWave;
    hipLaunchKernelGGL(k_batch_sys_sums, dim3((nsys + per - 1) / per), dim3(kBlock), 0, s,
                       nsys, A, eatom, out);
    return hipGetLastError();
}

}  // namespace parm
