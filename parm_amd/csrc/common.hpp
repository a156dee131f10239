// common.hpp -- shared device helpers for the ParM MD hot path on gfx950.
//
// Everything here is compiled with -ffp-contract=off: an expression rounds as
// written unless it calls fma() explicitly.  The integrators rely on that to
// reproduce the reference's double-precision update bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace parm {

constexpr int kBlock = 256;   // threads per workgroup for per-atom kernels
constexpr int kWave = 64;     // CDNA wavefront

// ---------------------------------------------------------------- vectors
template <int D>
struct Vec {
    double c[D];
};

// Position storage: 3D positions live in a double4 (one 32-byte load per
// neighbour gather, the 4th lane unused); 2D positions in a double2.
template <int D>
struct PosStore;
template <>
struct PosStore<3> {
    using T = double4;
    __host__ __device__ static inline Vec<3> load(const T &p) {
        return Vec<3>{{p.x, p.y, p.z}};
    }
    __host__ __device__ static inline T make(const Vec<3> &v) {
        return T{v.c[0], v.c[1], v.c[2], 0.0};
    }
};
template <>
struct PosStore<2> {
    using T = double2;
    __host__ __device__ static inline Vec<2> load(const T &p) {
        return Vec<2>{{p.x, p.y}};
    }
    __host__ __device__ static inline T make(const Vec<2> &v) {
        return T{v.c[0], v.c[1]};
    }
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Renumbering them so that XCD k runs one contiguous range of logical
// workgroups keeps spatially adjacent atoms (adjacent slots) on one L2, so
// the neighbour positions they share are fetched into one L2, not eight.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds, k = b / kXcds;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// log2 of the lanes per atom (a power of two: 1, 2, 4 or 8), so that list
// entry positions take shifts and masks instead of integer divisions.
__device__ __forceinline__ int tpa_shift(int tpa) { return 31 - __clz(tpa); }

// Position of entry q of slot i in the 32-bit (untiled) list:
// ((q / tpa) * npad + i) * tpa + q % tpa.
__device__ __forceinline__ size_t untiled_entry(int q, int i, int tpa, int npad) {
    const int ts = tpa_shift(tpa);
    return (((size_t)(q >> ts) * npad + i) << ts) + (q & (tpa - 1));
}

// Box: edge lengths, their halves and inverses.
struct Box {
    double L[3];
    double invL[3];
    double halfL[3];
};

// |r|^2 summed as (x*x + y*y) + z*z, the order the reference's Eigen
// Vec::squaredNorm uses for a 3-vector of doubles.
template <int D>
__device__ __forceinline__ double sqnorm(const Vec<D> &r) {
    if constexpr (D == 3)
        return (r.c[0] * r.c[0] + r.c[1] * r.c[1]) + r.c[2] * r.c[2];
    else
        return r.c[0] * r.c[0] + r.c[1] * r.c[1];
}

// Minimum image of one component, equal to IEEE remainder(dx, L) bit for bit
// (OriginBox::diff, src/box.hpp:68-78,103).
//
// remainder(dx, L) = dx - n* L with n* the integer nearest dx/L (ties to
// even), and that value is always representable, so fma(-n*, L, dx) returns
// it exactly.  n = rint(dx * (1/L)) is within one of n* whenever |dx/L| <
// 2^50.  For any other integer m, |dx - m L| >= L/2 exactly, and rounding is
// monotone with L/2 representable, so |fma(-m, L, dx)| >= L/2 too.  Hence:
// if |fma(-n, L, dx)| < L/2, n = n*.  Otherwise n* is the one of n-1, n,
// n+1 whose result is below L/2, or, on an exact tie (two results of
// magnitude L/2), the even one.
__device__ __forceinline__ double min_image1(double dx, double L, double invL,
                                             double halfL) {
    const double n = rint(dx * invL);
    double r = fma(-n, L, dx);
    if (__builtin_expect(!(fabs(r) < halfL), 0)) {
        double best = r, bestn = n;
        bool found = false;
        for (int d = -1; d <= 1; d++) {
            const double m = n + (double)d;
            const double rm = fma(-m, L, dx);
            const double am = fabs(rm);
            if (am < halfL) {
                best = rm;
                bestn = m;
                found = true;
            } else if (!found && am == halfL && rint(m * 0.5) == m * 0.5) {
                best = rm;
                bestn = m;
            }
        }
        (void)bestn;
        r = best;
    }
    return r;
}

template <int D>
__device__ __forceinline__ Vec<D> diff(const Vec<D> &a, const Vec<D> &b,
                                       const Box &box) {
    Vec<D> r;
#pragma unroll
    for (int k = 0; k < D; k++)
        r.c[k] = min_image1(a.c[k] - b.c[k], box.L[k], box.invL[k],
                            box.halfL[k]);
    return r;
}

// NeighborList pair predicate (src/trackers.cpp:63-67), bit for bit:
//   r = diff(x_i, x_j), d2 = (x^2 + y^2) + z^2,
//   in list  <=>  sqrt(d2) < (d_i + d_j)/2 + skin   (correctly rounded sqrt).
// sqrt is only evaluated when d2 is within 1e-14 (relative) of c^2; outside
// that band the comparison with c^2 decides it with margin to spare.
template <int D>
__device__ __forceinline__ bool in_list(const Vec<D> &xi, const Vec<D> &xj, double di,
                                        double dj, const Box &box, double skin) {
    const Vec<D> r = diff<D>(xi, xj, box);
    const double d2 = sqnorm<D>(r);
    const double c = (di + dj) / 2 + skin;
    const double c2 = c * c;
    if (d2 < c2 * (1.0 - 1e-14)) return true;
    if (d2 > c2 * (1.0 + 1e-14)) return false;
    return sqrt(d2) < c;
}


// 1/x with one Newton-Raphson step on the hardware estimate: the estimate's
// relative error (measured on gfx950: at most 4.6e-8 over 2^24 arguments
// spanning 20 binades) is squared by the step, leaving 1/x within 2.2e-15
// relative (11 ulp); two steps (rcp64) round correctly on the same sample.
__device__ __forceinline__ double rcp64_1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// 1/x to ~1 ulp: hardware estimate and two Newton-Raphson steps.
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    return r;
}

// ------------------------------------------------------------ reductions
// DPP lane moves (VALU operand modifiers: no trip through the LDS crossbar).
// A 64-bit value moves as its two 32-bit halves.  The patterns used here
// read a valid source lane for every lane, so every lane is written.
constexpr int kDppQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kDppQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]: lane ^ 2
constexpr int kDppRowHalfMirror = 0x141; // lane 7 - l within each 8 lanes
constexpr int kDppRowMirror = 0x140;     // lane 15 - l within each 16 lanes

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long x = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(
        (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long x = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 4 lanes of each quad (every lane of the quad gets it), in the
// order of the xor-1 then xor-2 butterfly: (v0 + v1) + (v2 + v3).
__device__ __forceinline__ double quad_sum(double v) {
    v += dpp_f64<kDppQuadXor1>(v);
    return v + dpp_f64<kDppQuadXor2>(v);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Neumaier-compensated running sum (value + compensation).
struct KSum {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        double t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

// Compensated pair (sum, error) combined with Knuth's two-sum: used to merge
// Neumaier partial sums in a fixed tree.
struct CSum {
    double s = 0.0, c = 0.0;
};
__device__ __forceinline__ CSum csum_merge(CSum a, CSum b) {
    const double t = a.s + b.s;
    const double bp = t - a.s;
    const double err = (a.s - (t - bp)) + (b.s - bp);
    return CSum{t, a.c + b.c + err};
}

// Top-2 of non-negative values (ties kept: {d, d} has sum 2d).
struct Top2 {
    double a = 0.0, b = 0.0;  // a >= b
    __device__ __forceinline__ void add(double d) {
        if (d > a) {
            b = a;
            a = d;
        } else if (d > b) {
            b = d;
        }
    }
    __device__ __forceinline__ void merge(const Top2 &o) {
        add(o.a);
        add(o.b);
    }
};

// (a, b) := the top 2 of {a, b, a2, b2} (a >= b, a2 >= b2; no NaN: Top2
// never holds one).
__device__ __forceinline__ void top2_merge(double &a, double &b, double a2, double b2) {
    const double lo = fmin(a, a2);
    a = fmax(a, a2);
    b = fmax(lo, fmax(b, b2));
}

template <int CTRL>
__device__ __forceinline__ void top2_dpp_step(Top2 &t) {
    const double a2 = dpp_f64<CTRL>(t.a), b2 = dpp_f64<CTRL>(t.b);
    top2_merge(t.a, t.b, a2, b2);
}

// Top 2 over the wavefront (all lanes active), returned to every lane: DPP
// steps leave each row of 16 lanes with its row's top 2, then the four rows
// are merged from lanes 0, 16, 32, 48.  The top 2 of a multiset does not
// depend on the merge order.
__device__ __forceinline__ Top2 wave_top2(Top2 t) {
    top2_dpp_step<kDppQuadXor1>(t);
    top2_dpp_step<kDppQuadXor2>(t);
    top2_dpp_step<kDppRowHalfMirror>(t);
    top2_dpp_step<kDppRowMirror>(t);
    Top2 r;
    r.a = readlane_f64(t.a, 0);
    r.b = readlane_f64(t.b, 0);
#pragma unroll
    for (int row = 1; row < kWave / 16; row++)
        top2_merge(r.a, r.b, readlane_f64(t.a, 16 * row), readlane_f64(t.b, 16 * row));
    return r;
}

// Sum over the wavefront of small non-negative integer counts held as
// doubles (exact in any order), returned to every lane.
__device__ __forceinline__ double wave_count(double v) {
    v = quad_sum(v);
    v += dpp_f64<kDppRowHalfMirror>(v);
    v += dpp_f64<kDppRowMirror>(v);
    double r = readlane_f64(v, 0);
#pragma unroll
    for (int row = 1; row < kWave / 16; row++) r += readlane_f64(v, 16 * row);
    return r;
}

// Flag word stored by the reduction kernels in device memory and in a pinned
// host mirror.
enum : uint32_t {
    kFlagRebuild = 1u,      // skin criterion met: rebuild before the next forces
    kFlagNonFinite = 2u,    // a displacement was NaN/Inf
    kFlagOverflow = 4u,     // neighbour capacity exceeded during a build
    kFlagReplan = 8u,       // a speculative tiled build met a tile its bounds do not fit
};

}  // namespace parm
