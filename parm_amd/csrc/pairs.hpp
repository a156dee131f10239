// pairs.hpp -- per-pair force / energy math and the per-atom integrator
// updates shared by the single-system (step.hip) and batched (batch.hip)
// kernels.
//
// Reference behaviour restated (paths relative to the reference root):
//   LJAttractRepulsePair::forces/energy  src/interaction.hpp:1270-1298
//   RepulsionPair::forces/energy         src/interaction.hpp:1538-1556
//   CollectionVerlet::timestep           src/collection.cpp:442-469
//   CollectionDamped::timestep           src/collection.cpp:356-381
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace parm {

// ------------------------------------------------------------ pair math
struct PairAcc {
    double e = 0.0, w = 0.0, s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double contacts = 0.0, overlaps = 0.0;
};

// LJAttractRepulsePair's force factor fs (f = fs r) from d2 = |r|^2, with
// rsix = (sigma / r)^6 and fmagR = 12 eps rsix (rsix - 1) (the energy and
// virial terms) as by-products: one expression for every FP64 caller, so
// they all round alike.  1/d2 from one Newton step (rcp64_1: within 2.2e-15
// relative), which puts fs within ~2e-14 of its terms' scale, an order of
// magnitude inside the FP64 force bar's 1e-13 per term.
__device__ __forceinline__ double lj64_fs(const PairParam &pp, double d2, double &rsix,
                                          double &fmagR) {
    const double idsq = rcp64_1(d2);
    const double ir = pp.sig2 * idsq;  // sigma^2 / r^2
    rsix = ir * ir * ir;               // sigma^6 / r^6
    fmagR = pp.p6 * rsix * (rsix - 1.0);
    return fmagR * idsq;
}

// The same in FP32 (mixed precision).
__device__ __forceinline__ float lj32_fs(const PairParam &pp, float d2, float &rsix,
                                         float &fmagR) {
    const float idsq = __builtin_amdgcn_rcpf(d2);
    const float ir = pp.sig2_f * idsq;
    rsix = ir * ir * ir;
    fmagR = pp.p6_f * rsix * (rsix - 1.0f);
    return fmagR * idsq;
}

// FP64 LJAttractRepulsePair on one visit; returns false if outside cutoff.
// Cutoff: the reference skips when d2 / sig^2 > cut^2; lim_lo / lim_hi bracket
// that boundary so the division only runs within 1e-14 of it.
template <int D, bool ENERGY>
__device__ __forceinline__ void lj64(const PairParam &pp, const Vec<D> &r,
                                     double (&f)[D], PairAcc &acc) {
    const double d2 = sqnorm<D>(r);
    bool in = d2 < pp.lim_lo ||
              (!(d2 > pp.lim_hi) && !(d2 / pp.sig2 > pp.cut2));
    if (!in) return;
    double rsix, fmagR;
    const double fs = lj64_fs(pp, d2, rsix, fmagR);
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = fma(fs, r.c[k], f[k]);
    if (ENERGY) {
        const double mid = 1.0 - rsix;
        const double e = pp.eps * (mid * mid) - pp.p5;
        acc.e += e;
        acc.w += fmagR;
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int b = 0; b < D; b++) acc.s[a * D + b] += r.c[a] * (r.c[b] * fs);
        acc.contacts += (e != 0.0) ? 1.0 : 0.0;
        acc.overlaps += (e > 0.0) ? 1.0 : 0.0;
    }
}

// FP64 RepulsionPair: same operations as the reference, so each pair force
// is bit-identical to the reference's for n == 2 (pow(x,1) = x, pow(x,2) = x*x).
template <int D, bool ENERGY, bool HARM>
__device__ __forceinline__ void rep64(const PairParam &pp, const Vec<D> &r,
                                      double (&f)[D], PairAcc &acc) {
    const double d2 = sqnorm<D>(r);
    if (d2 > pp.sig2) return;
    const double R = sqrt(d2);
    const double x = 1.0 - (R / pp.sig);
    const double px = HARM ? x : pow(x, pp.p6);
    const double fs = pp.eps * px / pp.sig / R;
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = fma(r.c[k], fs, f[k]);
    if (ENERGY) {
        const double e = pp.eps * (HARM ? x * x : pow(x, pp.p5)) / pp.p5;
        acc.e += e;
        double rf = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) rf += r.c[k] * (r.c[k] * fs);
        acc.w += rf;
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int b = 0; b < D; b++) acc.s[a * D + b] += r.c[a] * (r.c[b] * fs);
        acc.contacts += (e != 0.0) ? 1.0 : 0.0;
        acc.overlaps += (e > 0.0) ? 1.0 : 0.0;
    }
}

// Mixed precision: r^2, r^-6 and the force are evaluated in FP32 from an FP32
// displacement rf.  The in/out decision near the cutoff must not flip with
// FP32 rounding (the LJ force is not shifted to zero there): within 1e-5 of
// the cutoff it is taken in FP64 exactly as the FP64 path takes it, on the
// d2 returned by exact_d2() (the FP64 minimum-image displacement).
template <int D, bool ENERGY, typename Exact>
__device__ __forceinline__ void lj32f(const PairParam &pp, const float (&rf)[D],
                                      const Exact &exact_d2, float (&f)[D], PairAcc &acc) {
    float d2 = rf[0] * rf[0];
#pragma unroll
    for (int k = 1; k < D; k++) d2 = fmaf(rf[k], rf[k], d2);
    if (d2 > pp.lim_f * 0.99999f) {
        if (d2 > pp.lim_f * 1.00001f) return;
        const double d2d = exact_d2();
        if (!(d2d < pp.lim_lo || (!(d2d > pp.lim_hi) && !(d2d / pp.sig2 > pp.cut2))))
            return;
    }
    float rsix, fmagR;
    const float fs = lj32_fs(pp, d2, rsix, fmagR);
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = fmaf(fs, rf[k], f[k]);
    if (ENERGY) {
        const float mid = 1.0f - rsix;
        const double e = (double)(pp.eps_f * (mid * mid) - pp.p5_f);
        acc.e += e;
        acc.w += (double)fmagR;
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int b = 0; b < D; b++)
                acc.s[a * D + b] += (double)(rf[a] * (rf[b] * fs));
        acc.contacts += (e != 0.0) ? 1.0 : 0.0;
        acc.overlaps += (e > 0.0) ? 1.0 : 0.0;
    }
}

template <int D, bool ENERGY, bool HARM, typename Exact>
__device__ __forceinline__ void rep32f(const PairParam &pp, const float (&rf)[D],
                                       const Exact &exact_d2, float (&f)[D], PairAcc &acc) {
    float d2 = rf[0] * rf[0];
#pragma unroll
    for (int k = 1; k < D; k++) d2 = fmaf(rf[k], rf[k], d2);
    if (d2 > pp.sig2_f * 0.99999f) {  // contact decided in FP64 near sigma
        if (d2 > pp.sig2_f * 1.00001f) return;
        if (exact_d2() > pp.sig2) return;
    }
    const float R = __builtin_sqrtf(d2);
    const float x = fmaxf(1.0f - R * pp.isig_f, 0.0f);
    const float px = HARM ? x : powf(x, pp.p6_f);
    const float fs = pp.eps_f * px * pp.isig_f * __builtin_amdgcn_rcpf(R);
#pragma unroll
    for (int k = 0; k < D; k++) f[k] = fmaf(rf[k], fs, f[k]);
    if (ENERGY) {
        const double e =
            (double)(pp.eps_f * (HARM ? x * x : powf(x, pp.p5_f)) / pp.p5_f);
        acc.e += e;
        float rfs = 0.0f;
#pragma unroll
        for (int k = 0; k < D; k++) rfs = fmaf(rf[k], rf[k] * fs, rfs);
        acc.w += (double)rfs;
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int b = 0; b < D; b++)
                acc.s[a * D + b] += (double)(rf[a] * (rf[b] * fs));
        acc.contacts += (e != 0.0) ? 1.0 : 0.0;
        acc.overlaps += (e > 0.0) ? 1.0 : 0.0;
    }
}

// Untiled mixed path: the displacement is formed in FP64 (exact minimum
// image) and rounded to FP32.
template <int D, bool ENERGY>
__device__ __forceinline__ void lj32(const PairParam &pp, const Vec<D> &r,
                                     float (&f)[D], PairAcc &acc) {
    float rf[D];
#pragma unroll
    for (int k = 0; k < D; k++) rf[k] = (float)r.c[k];
    lj32f<D, ENERGY>(pp, rf, [&]() { return sqnorm<D>(r); }, f, acc);
}

template <int D, bool ENERGY, bool HARM>
__device__ __forceinline__ void rep32(const PairParam &pp, const Vec<D> &r,
                                      float (&f)[D], PairAcc &acc) {
    float rf[D];
#pragma unroll
    for (int k = 0; k < D; k++) rf[k] = (float)r.c[k];
    rep32f<D, ENERGY, HARM>(pp, rf, [&]() { return sqnorm<D>(r); }, f, acc);
}

// ------------------------------------------------------------ step kernels
__device__ __forceinline__ bool is_fixed(double m) { return m <= 0 || isinf(m); }

// First half of a step for one atom: the reference's first loop.
template <int D, bool DAMPED>
__device__ __forceinline__ void drift_atom(Vec<D> &x, double (&v)[D],
                                           const double (&a)[D], bool fixed,
                                           const StepCoef &c) {
    if (fixed) {
#pragma unroll
        for (int k = 0; k < D; k++) v[k] = 0.0;
        return;
    }
#pragma unroll
    for (int k = 0; k < D; k++) {
        if (!DAMPED) {
            // x += v*dt + a*(dt*dt/2) ; v += a*(dt/2)
            x.c[k] = x.c[k] + (v[k] * c.k0 + a[k] * c.k1);
            v[k] = v[k] + a[k] * c.k2;
        } else {
            // x += v*(c1 dt) + a*(c2 dt dt) ; v = v*c0 + a*(dt (c1 - c2))
            x.c[k] = x.c[k] + (v[k] * c.k0 + a[k] * c.k1);
            v[k] = v[k] * c.k3 + a[k] * c.k2;
        }
    }
}

// One pair visit: dispatch on kind and precision.
template <int D, int KIND, int PREC, bool ENERGY, bool HARM>
__device__ __forceinline__ void visit(const PairParam &pp, const Vec<D> &r,
                                      double (&f64)[D], float (&f32)[D],
                                      PairAcc &acc) {
    if (PREC == kFP64) {
        if (KIND == kPairLJ)
            lj64<D, ENERGY>(pp, r, f64, acc);
        else
            rep64<D, ENERGY, HARM>(pp, r, f64, acc);
    } else {
        if (KIND == kPairLJ)
            lj32<D, ENERGY>(pp, r, f32, acc);
        else
            rep32<D, ENERGY, HARM>(pp, r, f32, acc);
    }
}

}  // namespace parm
