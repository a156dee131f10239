// stepcore.hpp -- pieces shared by the force kernels of step.hip (untiled
// list) and step_tiled.hip (tiled list): per-workgroup reduction partials,
// the skin test of one atom, and the per-atom epilogue that applies the
// integrator stage fused into a force launch.
//
// Reference behaviour restated (paths relative to the reference root):
//   Collection::set_forces               src/collection.cpp:159-179
//   CollectionVerlet::timestep           src/collection.cpp:442-469
//   CollectionDamped::timestep           src/collection.cpp:356-381
//   NeighborList::update_list skin test  src/trackers.cpp:23-53
#pragma once

#include "common.hpp"
#include "kernels.hpp"
#include "pairs.hpp"

namespace parm {

// ------------------------------------------------------ reductions
// A kernel that reduces (sums of energy / virial / stress / counts, and the
// skin test's top-2 and non-finite count) writes one partial per workgroup
// with plain stores; k_finalize, launched right after it on the same stream,
// reduces the partials in a fixed order.  (A last-workgroup reduction inside
// the launch would need a device-scope release per workgroup, which on a
// multi-XCD part writes back the XCD's L2: far dearer than one more launch.)  No fences or contended atomics are
// needed inside the hot kernel, and the result does not depend on scheduling.
constexpr int kSumU = 0, kSumW = 1, kSumS = 2, kSumContacts = 11,
              kSumOverlaps = 12, kTop1 = 13, kTop2 = 14, kNonFinite = 15;

// Flag word of a finished skin test: top-2 displacements tt, nft non-finite
// displacements.  NeighborList::update_list (src/trackers.cpp:23-53) checks
// bigdist + biggestdist >= skin only after one of the two grew, which needs a
// displacement > 0: when nothing moved it never rebuilds.
__device__ __forceinline__ uint32_t skin_flag(const Top2 &tt, double nft, double skin) {
    uint32_t fl = 0;
    if (nft > 0) fl |= kFlagNonFinite;
    if (tt.a > 0 && tt.b + tt.a >= skin) fl |= kFlagRebuild;
    return fl;
}

// SKIN_W0: only the first wavefront holds skin-test values (the others'
// are the neutral ones and skip their wavefront reduction).
template <bool SUMS, bool SKIN, int NW, bool SKIN_W0 = false>
__device__ __forceinline__ void block_partial(double (&v)[kRedWidth], Top2 t,
                                              const ReduceBuf &R) {
    __shared__ double sh[NW][kRedWidth];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (SUMS) {
#pragma unroll
        for (int k = 0; k < kTop1; k++) v[k] = wave_sum(v[k]);
    }
    if (SKIN && (!SKIN_W0 || w == 0)) {
        t = wave_top2(t);
        v[kNonFinite] = wave_count(v[kNonFinite]);
    }
    if (lane == 0) {
        if (SUMS) {
#pragma unroll
            for (int k = 0; k < kTop1; k++) sh[w][k] = v[k];
        }
        if (SKIN) {
            sh[w][kTop1] = t.a;
            sh[w][kTop2] = t.b;
            sh[w][kNonFinite] = v[kNonFinite];
        }
    }
    __syncthreads();
    const int k = threadIdx.x;
    // field-major partials: field k of workgroup b at part[k * pstride + b]
    double *dst = R.part + blockIdx.x;
    const size_t ps = (size_t)R.pstride;
    if (SUMS && k < kTop1) {
        double p = 0.0;
        for (int q = 0; q < NW; q++) p += sh[q][k];
        dst[k * ps] = p;
    }
    if (SKIN && k == kTop1) {
        Top2 bt;
        double nf = 0.0;
        for (int q = 0; q < NW; q++) {
            Top2 o;
            o.a = sh[q][kTop1];
            o.b = sh[q][kTop2];
            bt.merge(o);
            nf += sh[q][kNonFinite];
        }
        dst[kTop1 * ps] = bt.a;
        dst[kTop2 * ps] = bt.b;
        dst[kNonFinite * ps] = nf;
    }
}

template <int D>
__device__ __forceinline__ void skin_atom(const Vec<D> &x, const void *xref, int i,
                                          Top2 &t, double &nonfinite) {
    const Vec<D> x0 =
        PosStore<D>::load(static_cast<const typename PosStore<D>::T *>(xref)[i]);
    Vec<D> d;
#pragma unroll
    for (int k = 0; k < D; k++) d.c[k] = x.c[k] - x0.c[k];
    const double dist = sqrt(sqnorm<D>(d));
    t.add(dist);
    if (!isfinite(dist)) nonfinite += 1.0;
}

// Per-atom inputs of the stage's update (loaded ahead by the tiled kernel).
template <int D>
struct AtomIn {
    double m;
    double v[D];
    Vec<D> x, xref;
    bool member;
};

template <int D, int STAGE, bool LOADX = true>
__device__ __forceinline__ void atom_load(const AtomArrays &A, int i, AtomIn<D> &in) {
    using PT = typename PosStore<D>::T;
    constexpr bool DRIFT =
        (STAGE == kStageVerletKickDrift || STAGE == kStageDampedKickDrift);
    if (STAGE == kStageSums || STAGE == kStageForceOnly || STAGE == kStageForceAdd) return;
    in.m = A.m[i];
    if (STAGE != kStageForceA) {
#pragma unroll
        for (int k = 0; k < D; k++) in.v[k] = A.v[k][i];
    }
    if (DRIFT) {
        if (LOADX) in.x = PosStore<D>::load(static_cast<const PT *>(A.pos_in)[i]);
        in.member = A.rank[i] >= 0;
        in.xref = PosStore<D>::load(static_cast<const PT *>(A.xref)[i]);
    }
}

// The stage's writes for atom i with total pair force fi and inputs `in`;
// the skin test of the drifted position feeds top / nonfinite.
template <int D, int STAGE>
__device__ __forceinline__ void atom_store(const AtomArrays &A, int i, const double (&fi)[D],
                                           const AtomIn<D> &in, const StepCoef &c, Top2 &top,
                                           double &nonfinite) {
    using PT = typename PosStore<D>::T;
    constexpr bool DRIFT =
        (STAGE == kStageVerletKickDrift || STAGE == kStageDampedKickDrift);
    constexpr bool DAMPED =
        (STAGE == kStageDampedKick || STAGE == kStageDampedKickDrift);
    if (STAGE == kStageSums) return;
    // The fused kick + drift stages run inside a step loop whose last step is
    // a plain kick launch: nothing reads f or a in between (the drift uses
    // the a it has just computed), so they are stored by the last step only.
    if constexpr (!DRIFT) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            if (STAGE == kStageForceAdd)
                A.f[k][i] += fi[k];
            else
                A.f[k][i] = fi[k];
        }
    }
    if (STAGE == kStageForceOnly || STAGE == kStageForceAdd) return;
    const bool fixed = is_fixed(in.m);
    double a[D];
#pragma unroll
    for (int k = 0; k < D; k++) a[k] = fixed ? 0.0 : fi[k] / in.m;
    if constexpr (!DRIFT) {
#pragma unroll
        for (int k = 0; k < D; k++) A.a[k][i] = a[k];
    }
    if (STAGE == kStageForceA) return;
    double v[D];
#pragma unroll
    for (int k = 0; k < D; k++) v[k] = in.v[k];
    if (!fixed) {
#pragma unroll
        for (int k = 0; k < D; k++) v[k] = v[k] + a[k] * (DAMPED ? c.k4 : c.k2);
    }
    if (DRIFT) {
        Vec<D> xn = in.x;
        drift_atom<D, DAMPED>(xn, v, a, fixed, c);
        static_cast<PT *>(A.pos_out)[i] = PosStore<D>::make(xn);
        if (in.member) {
            Vec<D> d;
#pragma unroll
            for (int k = 0; k < D; k++) d.c[k] = xn.c[k] - in.xref.c[k];
            const double dist = sqrt(sqnorm<D>(d));
            top.add(dist);
            if (!isfinite(dist)) nonfinite += 1.0;
        }
    }
#pragma unroll
    for (int k = 0; k < D; k++) A.v[k][i] = v[k];
}

// This workgroup's reduction partials (all threads call it).  SKIN_W0: the
// stage's writes (and so the skin test) were done by the first wavefront.
template <int D, int STAGE, bool ENERGY, int NW, bool SKIN_W0 = false>
__device__ __forceinline__ void group_partials(const PairAcc &acc, Top2 top, double nonfinite,
                                               const ReduceBuf &R) {
    constexpr bool DRIFT =
        (STAGE == kStageVerletKickDrift || STAGE == kStageDampedKickDrift);
    if (ENERGY || DRIFT) {
        double v[kRedWidth];
#pragma unroll
        for (int k = 0; k < kRedWidth; k++) v[k] = 0.0;
        if (ENERGY) {
            v[kSumU] = 0.5 * acc.e;
            v[kSumW] = 0.5 * acc.w;
#pragma unroll
            for (int k = 0; k < D * D; k++) v[kSumS + k] = 0.5 * acc.s[k];
            v[kSumContacts] = acc.contacts;
            v[kSumOverlaps] = acc.overlaps;
        }
        v[kNonFinite] = nonfinite;
        block_partial<ENERGY, DRIFT, NW, SKIN_W0>(v, top, R);
    }
}

}  // namespace parm
