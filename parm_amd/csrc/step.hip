// step.hip -- pair forces over the device Verlet list, fused with the
// velocity-Verlet / damped-Verlet update and the skin test.
//
// Reference behaviour restated (paths relative to the reference root):
//   NListed<A,P>::set_forces / energy / pressure / stress
//                                   src/interaction.hpp:2153-2291
//   LJAttractRepulsePair::forces/energy  src/interaction.hpp:1270-1298
//   RepulsionPair::forces/energy         src/interaction.hpp:1538-1556
//   Collection::set_forces               src/collection.cpp:159-179
//   CollectionVerlet::timestep           src/collection.cpp:442-469
//   CollectionDamped::timestep           src/collection.cpp:356-381
//   NeighborList::update_list skin test  src/trackers.cpp:23-53
//
// The device list is a FULL list (each pair appears in both atoms' rows), so
// each thread gathers the force on its own atom without atomics and in a
// fixed order.  The per-pair force from j on i uses r = diff(x_i, x_j); the
// visit from j computes -r bit for bit, so the two visits are exact negatives
// and only the summation order differs from the reference's half-list
// scatter.  Energies, virials and contact counts are halved.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"
#include "pairs.hpp"
#include "stepcore.hpp"

namespace parm {

// One workgroup of kFinal threads: thread t reduces partials t, t + kFinal,
// ... of each field (Neumaier; the loads of a field are coalesced across the
// workgroup), then a fixed butterfly per wave and a fixed combine over waves
// (compensated two-sum), so the result is deterministic.
constexpr int kFinal = 1024;

template <bool SUMS, bool SKIN>
__global__ __launch_bounds__(kFinal) void k_finalize(int nparts, ReduceBuf R) {
    const int tid = threadIdx.x;
    const int lane = tid % kWave, w = tid / kWave;
    constexpr int NW = kFinal / kWave;
    const size_t ps = (size_t)R.pstride;
    if constexpr (SUMS) {
        CSum cs[kTop1];
#pragma unroll 4
        for (int b = tid; b < nparts; b += kFinal) {
#pragma unroll
            for (int k = 0; k < kTop1; k++) {
                KSum q{cs[k].s, cs[k].c};
                q.add(R.part[k * ps + b]);
                cs[k] = CSum{q.s, q.c};
            }
        }
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < kTop1; k++) {
                CSum o{__shfl_xor(cs[k].s, off), __shfl_xor(cs[k].c, off)};
                cs[k] = csum_merge(cs[k], o);
            }
        }
        __shared__ double fin[NW][2 * kTop1];
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < kTop1; k++) {
                fin[w][2 * k] = cs[k].s;
                fin[w][2 * k + 1] = cs[k].c;
            }
        }
        __syncthreads();
        if (tid < kTop1) {
            CSum tot;
            for (int q = 0; q < NW; q++)
                tot = csum_merge(tot, CSum{fin[q][2 * tid], fin[q][2 * tid + 1]});
            R.out[tid] = tot.s + tot.c;
        }
    }
    if constexpr (SKIN) {
        Top2 bt;
        double nf = 0.0;
        // deep unroll: a partial array of ~10^4 workgroups is read in one or
        // two rounds of loads in flight
#pragma unroll 16
        for (int b = tid; b < nparts; b += kFinal) {
            Top2 o;
            o.a = R.part[kTop1 * ps + b];
            o.b = R.part[kTop2 * ps + b];
            bt.merge(o);
            nf += R.part[kNonFinite * ps + b];
        }
        bt = wave_top2(bt);
        nf = wave_count(nf);
        __shared__ double fin2[NW][3];
        if (lane == 0) {
            fin2[w][0] = bt.a;
            fin2[w][1] = bt.b;
            fin2[w][2] = nf;
        }
        __syncthreads();
        if (tid == 0) {
            Top2 tt;
            double nft = 0.0;
            for (int q = 0; q < NW; q++) {
                Top2 o;
                o.a = fin2[q][0];
                o.b = fin2[q][1];
                tt.merge(o);
                nft += fin2[q][2];
            }
            R.out[kTop1] = tt.a;
            R.out[kTop2] = tt.b;
            R.out[kNonFinite] = nft;
            const uint32_t fl = skin_flag(tt, nft, R.skin);
            *R.flag = fl;
            if (R.hflag)
                __hip_atomic_store(R.hflag, fl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// The skin test's reduction over up to kFinBlocks workgroups of kFinThreads:
// workgroup g merges partials g, g + G, ... (top-2 and non-finite count; the
// merge does not depend on order), stores its result, and the last
// workgroup to finish (a counter: release and acquire at device scope, cheap
// here, where little is written) merges the G results and raises the flag.
// One workgroup over 16,384 partials (1M atoms) takes ~10 us; G of them, each
// reading a few hundred, a fraction of that.  The counter is returned to 0 by
// the last workgroup, so the next launch finds it cleared.
constexpr int kFinThreads = 256;

__device__ __forceinline__ void fin_block_top2(Top2 &t, double &nf) {
    constexpr int NW = kFinThreads / kWave;
    t = wave_top2(t);
    nf = wave_count(nf);
    __shared__ double sh[NW][3];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) {
        sh[w][0] = t.a;
        sh[w][1] = t.b;
        sh[w][2] = nf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Top2 tt;
        double n2 = 0.0;
        for (int q = 0; q < NW; q++) {
            Top2 o;
            o.a = sh[q][0];
            o.b = sh[q][1];
            tt.merge(o);
            n2 += sh[q][2];
        }
        t = tt;
        nf = n2;
    }
}

__global__ __launch_bounds__(kFinThreads) void k_finalize_skin(int nparts, ReduceBuf R) {
    const size_t ps = (size_t)R.pstride;
    const int G = (int)gridDim.x, g = (int)blockIdx.x;
    Top2 t;
    double nf = 0.0;
#pragma unroll 4
    for (int b = g * kFinThreads + (int)threadIdx.x; b < nparts; b += G * kFinThreads) {
        Top2 o;
        o.a = R.part[kTop1 * ps + b];
        o.b = R.part[kTop2 * ps + b];
        t.merge(o);
        nf += R.part[kNonFinite * ps + b];
    }
    fin_block_top2(t, nf);
    __shared__ int last;
    if (threadIdx.x == 0) {
        R.gpart[g] = t.a;
        R.gpart[kFinBlocks + g] = t.b;
        R.gpart[2 * kFinBlocks + g] = nf;
        __threadfence();
        last = atomicAdd(R.cnt, 1u) == (unsigned)(G - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    Top2 u;
    double nu = 0.0;
    if ((int)threadIdx.x < G) {
        u.a = R.gpart[threadIdx.x];
        u.b = R.gpart[kFinBlocks + threadIdx.x];
        nu = R.gpart[2 * kFinBlocks + threadIdx.x];
    }
    fin_block_top2(u, nu);
    if (threadIdx.x == 0) {
        __hip_atomic_store(R.cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        R.out[kTop1] = u.a;
        R.out[kTop2] = u.b;
        R.out[kNonFinite] = nu;
        const uint32_t fl = skin_flag(u, nu, R.skin);
        *R.flag = fl;
        if (R.hflag)
            __hip_atomic_store(R.hflag, fl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// below this many partials the one-workgroup k_finalize is the quicker (the
// counter's fences and the second merge cost more than they save: +6 % per
// step at 1000 atoms, +4 % at 32,768; -1.6 % at 1M atoms, 16,384 partials)
constexpr int kFinMultiMin = 8 * kFinal;

template <bool SUMS, bool SKIN>
static void launch_finalize(int nparts, const ReduceBuf &R, hipStream_t s) {
    if (!SUMS && SKIN && R.cnt && R.gpart && nparts >= kFinMultiMin) {
        const int G = std::max(1, std::min(kFinBlocks, (nparts + kFinThreads - 1) / kFinThreads));
        hipLaunchKernelGGL(k_finalize_skin, dim3(G), dim3(kFinThreads), 0, s, nparts, R);
        return;
    }
    hipLaunchKernelGGL((k_finalize<SUMS, SKIN>), dim3(1), dim3(kFinal), 0, s, nparts, R);
}

// TPA lanes share one atom: lane `sub` visits list entries sub, sub + TPA, ...
// The list is stored in groups of TPA entries per atom (entry q of slot i at
// ((q / TPA) * npad + i) * TPA + q % TPA), so the 64 lanes of a wave read
// 64 consecutive words per step of the loop.  Two entries per lane are in
// flight at a time (both list words, then both position gathers) to hide
// memory latency.  The TPA partial forces are summed with a fixed butterfly,
// so the result is deterministic.
template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE, bool ENERGY,
          int TPA>
__global__ __launch_bounds__(kBlock) void k_force(AtomArrays A, ListView Lv,
                                                  PairView P, Box box, StepCoef c,
                                                  ReduceBuf R) {
    using PT = typename PosStore<D>::T;
    const int gt = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int i = gt / TPA;
    const int sub = gt % TPA;
    const bool act = i < A.n;
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
    if (act) xi = PosStore<D>::load(pos[i]);
    // the integrator inputs of the atom, loaded now to arrive during the
    // pair loop (lane sub == 0 applies the stage)
    AtomIn<D> ain;
    if (act && sub == 0) atom_load<D, STAGE, false>(A, i, ain);
    const bool member = act && A.rank[i] >= 0;
    if (member && Lv.built) {
        // lane sub's entries m = 0 .. nl-1 sit at (m * npad + i) * TPA + sub
        const int nl = Lv.lcount[(size_t)i * TPA + sub];
        const size_t stride = (size_t)Lv.npad * TPA;
        const uint32_t *nb = Lv.nbr + (size_t)i * TPA + sub;
        const int tb = Lv.type_bits;
        const uint32_t tmask = (1u << tb) - 1u;
        const PairParam *row = P.tab;
        PairParam p0;
        if (SINGLE)
            p0 = P.p0;
        else
            row = P.tab + (size_t)A.type[i] * P.ntypes;
        int m = 0;
        size_t off = 0;
        for (; m + 1 < nl; m += 2, off += 2 * stride) {
            const uint32_t e0 = nb[off];
            const uint32_t e1 = nb[off + stride];
            const PT pj0 = pos[e0 >> tb];
            const PT pj1 = pos[e1 >> tb];
            const Vec<D> r0 = diff<D>(xi, PosStore<D>::load(pj0), box);
            visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : row[e0 & tmask], r0, fi,
                                               f32, acc);
            const Vec<D> r1 = diff<D>(xi, PosStore<D>::load(pj1), box);
            visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : row[e1 & tmask], r1, fi,
                                               f32, acc);
        }
        if (m < nl) {
            const uint32_t e0 = nb[off];
            const Vec<D> r0 = diff<D>(xi, PosStore<D>::load(pos[e0 >> tb]), box);
            visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : row[e0 & tmask], r0, fi,
                                               f32, acc);
        }
    }
    if (PREC != kFP64) {
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] = (double)f32[k];
    }
    // sum the TPA partial forces of the atom (all lanes take part)
#pragma unroll
    for (int o = 1; o < TPA; o <<= 1) {
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] += __shfl_xor(fi[k], o);
    }

    Top2 top;
    double nonfinite = 0.0;
    if (act && sub == 0) {
        ain.x = xi;
        atom_store<D, STAGE>(A, i, fi, ain, c, top, nonfinite);
    }
    group_partials<D, STAGE, ENERGY, kBlock / kWave>(acc, top, nonfinite, R);
}

// Standalone first half of a step (drift + half kick) and skin test.
template <int D, bool DAMPED>
__global__ __launch_bounds__(kBlock) void k_drift(AtomArrays A, StepCoef c,
                                                  ReduceBuf R) {
    using PT = typename PosStore<D>::T;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    Top2 top;
    double nonfinite = 0.0;
    if (i < A.n) {
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
        if (A.rank[i] >= 0) skin_atom<D>(x, A.xref, i, top, nonfinite);
    }
    double v[kRedWidth];
#pragma unroll
    for (int k = 0; k < kRedWidth; k++) v[k] = 0.0;
    v[kNonFinite] = nonfinite;
    block_partial<false, true, kBlock / kWave>(v, top, R);
}

// Skin test only (after a rebuild, or for update_list(false) on demand).
template <int D>
__global__ __launch_bounds__(kBlock) void k_skin(AtomArrays A, ReduceBuf R) {
    using PT = typename PosStore<D>::T;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    Top2 top;
    double nonfinite = 0.0;
    if (i < A.n && A.rank[i] >= 0) {
        const Vec<D> x = PosStore<D>::load(static_cast<const PT *>(A.pos_in)[i]);
        skin_atom<D>(x, A.xref, i, top, nonfinite);
    }
    double v[kRedWidth];
#pragma unroll
    for (int k = 0; k < kRedWidth; k++) v[k] = 0.0;
    v[kNonFinite] = nonfinite;
    block_partial<false, true, kBlock / kWave>(v, top, R);
}

// Atom-wise observables (AtomGroup::kinetic_energy / momentum / mass,
// src/box.cpp:239-260, 401-411) and force statistics.
template <int D, int MODE>
__global__ __launch_bounds__(kBlock) void k_atom_sums(AtomArrays A, Vec<D> ov,
                                                      ReduceBuf R) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[kRedWidth];
#pragma unroll
    for (int k = 0; k < kRedWidth; k++) v[k] = 0.0;
    Top2 top;
    if (i < A.n) {
        const double m = A.m[i];
        if (MODE == 0) {
            if (!(m == 0 || isinf(m))) {
                double c[D];
#pragma unroll
                for (int k = 0; k < D; k++) c[k] = A.v[k][i] - ov.c[k];
                double dd;
                if constexpr (D == 3)
                    dd = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
                else
                    dd = c[0] * c[0] + c[1] * c[1];
                v[0] = m / 2 * dd;
            }
            if (!is_fixed(m)) {
#pragma unroll
                for (int k = 0; k < D; k++) v[1 + k] = A.v[k][i] * m;
                v[4] = m;
            }
        } else {
            Vec<D> f;
#pragma unroll
            for (int k = 0; k < D; k++) f.c[k] = A.f[k][i];
            const double fn = sqrt(sqnorm<D>(f));
            v[0] = fn;
            top.add(fn);
        }
    }
    block_partial<true, MODE == 1, kBlock / kWave>(v, top, R);
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_upload(AtomArrays A, int what,
                                                   const int32_t *__restrict__ orig,
                                                   const double *__restrict__ src) {
    using PT = typename PosStore<D>::T;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= A.n) return;
    const size_t o = (size_t)orig[s] * D;
    if (what == 0) {
        Vec<D> x;
#pragma unroll
        for (int k = 0; k < D; k++) x.c[k] = src[o + k];
        static_cast<PT *>(A.pos_out)[s] = PosStore<D>::make(x);
        return;
    }
    double *const *dst = what == 1 ? A.v : (what == 2 ? A.a : A.f);
#pragma unroll
    for (int k = 0; k < D; k++) dst[k][s] = src[o + k];
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_download(AtomArrays A, int what,
                                                     const int32_t *__restrict__ orig,
                                                     double *__restrict__ dst) {
    using PT = typename PosStore<D>::T;
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= A.n) return;
    const size_t o = (size_t)orig[s] * D;
    if (what == 0) {
        const Vec<D> x = PosStore<D>::load(static_cast<const PT *>(A.pos_in)[s]);
#pragma unroll
        for (int k = 0; k < D; k++) dst[o + k] = x.c[k];
        return;
    }
    double *const *src = what == 1 ? A.v : (what == 2 ? A.a : A.f);
#pragma unroll
    for (int k = 0; k < D; k++) dst[o + k] = src[k][s];
}

// ======================================================== host launchers
int reduce_blocks(int n) { return n <= 0 ? 1 : (n + kBlock - 1) / kBlock; }
int force_blocks(int n, int tpa) { return reduce_blocks(n * tpa); }

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE, int TPA>
static void launch_force_tt(bool energy, const AtomArrays &A, const ListView &Lv,
                            const PairView &P, const Box &box, const StepCoef &c,
                            const ReduceBuf &R, hipStream_t s) {
    constexpr bool DRIFT =
        (STAGE == kStageVerletKickDrift || STAGE == kStageDampedKickDrift);
    // energies are accumulated by the sums stage only
    constexpr bool ENERGY = STAGE == kStageSums;
    (void)energy;
    const int nb = force_blocks(A.n, TPA);
    dim3 g(nb), b(kBlock);
    hipLaunchKernelGGL((k_force<D, KIND, PREC, SINGLE, HARM, STAGE, ENERGY, TPA>), g, b, 0,
                       s, A, Lv, P, box, c, R);
    if (R.done) (void)hipEventRecord(R.done, s);
    if (ENERGY || DRIFT) launch_finalize<ENERGY, DRIFT>(nb, R, s);
}

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE>
static void launch_force_t(bool energy, const AtomArrays &A, const ListView &Lv,
                           const PairView &P, const Box &box, const StepCoef &c,
                           const ReduceBuf &R, hipStream_t s) {
    switch (Lv.tpa) {
    case 1:
        launch_force_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 1>(energy, A, Lv, P, box, c,
                                                               R, s);
        break;
    case 2:
        launch_force_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 2>(energy, A, Lv, P, box, c,
                                                               R, s);
        break;
    case 8:
        launch_force_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 8>(energy, A, Lv, P, box, c,
                                                               R, s);
        break;
    default:
        launch_force_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 4>(energy, A, Lv, P, box, c,
                                                               R, s);
        break;
    }
}

template <int D, int KIND, int PREC, bool SINGLE, bool HARM>
static void launch_force_s(int stage, bool energy, const AtomArrays &A,
                           const ListView &Lv, const PairView &P, const Box &box,
                           const StepCoef &c, const ReduceBuf &R, hipStream_t s) {
    switch (stage) {
#define PARM_STAGE(S)                                                          \
    case S:                                                                    \
        launch_force_t<D, KIND, PREC, SINGLE, HARM, S>(energy, A, Lv, P, box,  \
                                                       c, R, s);               \
        break;
        PARM_STAGE(kStageForceOnly)
        PARM_STAGE(kStageForceA)
        PARM_STAGE(kStageVerletKick)
        PARM_STAGE(kStageVerletKickDrift)
        PARM_STAGE(kStageDampedKick)
        PARM_STAGE(kStageDampedKickDrift)
        PARM_STAGE(kStageSums)
        PARM_STAGE(kStageForceAdd)
#undef PARM_STAGE
    }
}

template <int D, int KIND, int PREC>
static void launch_force_k(int stage, bool energy, const AtomArrays &A,
                           const ListView &Lv, const PairView &P, const Box &box,
                           const StepCoef &c, const ReduceBuf &R, hipStream_t s) {
    if constexpr (KIND == kPairRepulsion) {
        if (P.harmonic) {
            if (P.single)
                launch_force_s<D, KIND, PREC, true, true>(stage, energy, A, Lv, P, box,
                                                          c, R, s);
            else
                launch_force_s<D, KIND, PREC, false, true>(stage, energy, A, Lv, P, box,
                                                           c, R, s);
            return;
        }
    }
    {
        if (P.single)
            launch_force_s<D, KIND, PREC, true, false>(stage, energy, A, Lv, P, box,
                                                       c, R, s);
        else
            launch_force_s<D, KIND, PREC, false, false>(stage, energy, A, Lv, P, box,
                                                        c, R, s);
    }
}

template <int D>
static void launch_force_d(int precision, int stage, bool energy, const AtomArrays &A,
                           const ListView &Lv, const PairView &P, const Box &box,
                           const StepCoef &c, const ReduceBuf &R, hipStream_t s) {
    if (P.kind == kPairRepulsion) {
        if (precision == kFP64)
            launch_force_k<D, kPairRepulsion, kFP64>(stage, energy, A, Lv, P, box, c,
                                                     R, s);
        else
            launch_force_k<D, kPairRepulsion, kMixed>(stage, energy, A, Lv, P, box,
                                                      c, R, s);
    } else {
        if (precision == kFP64)
            launch_force_k<D, kPairLJ, kFP64>(stage, energy, A, Lv, P, box, c, R, s);
        else
            launch_force_k<D, kPairLJ, kMixed>(stage, energy, A, Lv, P, box, c, R,
                                               s);
    }
}

hipError_t launch_finalize_rt(bool sums, bool skin, int nparts, const ReduceBuf &R,
                              hipStream_t s) {
    if (sums && skin)
        launch_finalize<true, true>(nparts, R, s);
    else if (sums)
        launch_finalize<true, false>(nparts, R, s);
    else if (skin)
        launch_finalize<false, true>(nparts, R, s);
    return hipGetLastError();
}

hipError_t launch_force(int ndim, int precision, int stage, bool energy,
                        const AtomArrays &A, const ListView &Lv, const PairView &P,
                        const Box &box, const StepCoef &c, const ReduceBuf &R,
                        hipStream_t s) {
    if (energy != (stage == kStageSums)) return hipErrorInvalidValue;
    if (A.n <= 0) return hipSuccess;
    if (Lv.tiled && Lv.built)
        return launch_force_tiled(ndim, precision, stage, energy, A, Lv, P, box, c, R, s);
    if (ndim == 3)
        launch_force_d<3>(precision, stage, energy, A, Lv, P, box, c, R, s);
    else
        launch_force_d<2>(precision, stage, energy, A, Lv, P, box, c, R, s);
    return hipGetLastError();
}

hipError_t launch_drift(int ndim, int stage, const AtomArrays &A, const StepCoef &c,
                        const ReduceBuf &R, hipStream_t s) {
    if (A.n <= 0) return hipSuccess;
    dim3 g(reduce_blocks(A.n)), b(kBlock);
    const bool damped = (stage == kStageDampedKick || stage == kStageDampedKickDrift);
    if (ndim == 3) {
        if (damped)
            hipLaunchKernelGGL((k_drift<3, true>), g, b, 0, s, A, c, R);
        else
            hipLaunchKernelGGL((k_drift<3, false>), g, b, 0, s, A, c, R);
    } else {
        if (damped)
            hipLaunchKernelGGL((k_drift<2, true>), g, b, 0, s, A, c, R);
        else
            hipLaunchKernelGGL((k_drift<2, false>), g, b, 0, s, A, c, R);
    }
    launch_finalize<false, true>((int)g.x, R, s);
    return hipGetLastError();
}

hipError_t launch_atom_sums(int ndim, int mode, const AtomArrays &A,
                            const double *originv, const ReduceBuf &R, hipStream_t s) {
    dim3 g(reduce_blocks(A.n)), b(kBlock);
    if (ndim == 3) {
        Vec<3> ov{{0, 0, 0}};
        if (originv)
            for (int k = 0; k < 3; k++) ov.c[k] = originv[k];
        if (mode == 0)
            hipLaunchKernelGGL((k_atom_sums<3, 0>), g, b, 0, s, A, ov, R);
        else
            hipLaunchKernelGGL((k_atom_sums<3, 1>), g, b, 0, s, A, ov, R);
    } else {
        Vec<2> ov{{0, 0}};
        if (originv)
            for (int k = 0; k < 2; k++) ov.c[k] = originv[k];
        if (mode == 0)
            hipLaunchKernelGGL((k_atom_sums<2, 0>), g, b, 0, s, A, ov, R);
        else
            hipLaunchKernelGGL((k_atom_sums<2, 1>), g, b, 0, s, A, ov, R);
    }
    if (mode == 0)
        launch_finalize<true, false>((int)g.x, R, s);
    else
        launch_finalize<true, true>((int)g.x, R, s);
    return hipGetLastError();
}

hipError_t launch_upload(int ndim, int what, const AtomArrays &A, const int32_t *orig,
                         const double *src, hipStream_t s) {
    if (A.n <= 0) return hipSuccess;
    dim3 g(reduce_blocks(A.n)), b(kBlock);
    if (ndim == 3)
        hipLaunchKernelGGL(k_upload<3>, g, b, 0, s, A, what, orig, src);
    else
        hipLaunchKernelGGL(k_upload<2>, g, b, 0, s, A, what, orig, src);
    return hipGetLastError();
}

hipError_t launch_download(int ndim, int what, const AtomArrays &A,
                           const int32_t *orig, double *dst, hipStream_t s) {
    if (A.n <= 0) return hipSuccess;
    dim3 g(reduce_blocks(A.n)), b(kBlock);
    if (ndim == 3)
        hipLaunchKernelGGL(k_download<3>, g, b, 0, s, A, what, orig, dst);
    else
        hipLaunchKernelGGL(k_download<2>, g, b, 0, s, A, what, orig, dst);
    return hipGetLastError();
}

hipError_t launch_skin(int ndim, const AtomArrays &A, const ReduceBuf &R,
                       hipStream_t s) {
    if (A.n <= 0) return hipSuccess;
    dim3 g(reduce_blocks(A.n)), b(kBlock);
    if (ndim == 3)
        hipLaunchKernelGGL((k_skin<3>), g, b, 0, s, A, R);
    else
        hipLaunchKernelGGL((k_skin<2>), g, b, 0, s, A, R);
    launch_finalize<false, true>((int)g.x, R, s);
    return hipGetLastError();
}

}  // namespace parm
