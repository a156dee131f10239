// nlcg.hip -- device vector operations of the pressure-controlled nonlinear
// conjugate-gradient minimiser (CollectionNLCG, src/collection.cpp:494-855),
// and the device side of the asynchronous state snapshots.
//
// The minimiser's scalar control flow (secant line search, Polak-Ribiere
// beta, restarts) stays on the host: each of its steps needs one of these
// O(N) passes over the device state, reduced with the deterministic
// compensated partial sums the force kernels use (stepcore.hpp), so a step
// costs a few launches and one read of four doubles.
//
// Reference behaviour restated (paths relative to the reference root):
//   CollectionNLCG::stepx           src/collection.cpp:608-625
//   CollectionNLCGFixedL::stepx     src/collection.cpp:855-870
//   fdotf / fdota / fdotv / vdotv   src/collection.cpp:637-683
//   kinetic_energy                  src/collection.cpp:582-595
//   set_forces (a, v from f)        src/collection.cpp:554-568
//   timestep (a = f, v = a + v b)   src/collection.cpp:805-828
//   descend                         src/collection.cpp:837-853
#include "common.hpp"
#include "kernels.hpp"
#include "nlcg.hpp"
#include "stepcore.hpp"

namespace parm {

// x <- x * Lfac + v * dx for every atom (fixed atoms included: their v is 0),
// Lfac per dimension (CollectionNLCGFixedL scales the first one only).
template <int D>
__global__ __launch_bounds__(kBlock) void k_nlcg_stepx(AtomArrays A, Vec<D> lfac, double dx) {
    using PT = typename PosStore<D>::T;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= A.n) return;
    PT *pos = static_cast<PT *>(A.pos_out);
    Vec<D> x = PosStore<D>::load(pos[i]);
#pragma unroll
    for (int k = 0; k < D; k++) {
        const double xs = x.c[k] * lfac.c[k];
        x.c[k] = xs + A.v[k][i] * dx;
    }
    pos[i] = PosStore<D>::make(x);
}

// Dot products in the reference's Eigen order for one atom.
template <int D>
__device__ __forceinline__ double dotv(const double (&p)[D], const double (&q)[D]) {
    if constexpr (D == 3)
        return (p[0] * q[0] + p[1] * q[1]) + p[2] * q[2];
    else
        return p[0] * q[0] + p[1] * q[1];
}

// One pass: the assignment `op` (NlcgOp), then, for atoms with a finite
// positive mass, the partial sums f.f, f.a, f.v, v.v (fields 0..3) and
// |v + x lfac|^2 (field 4).
template <int D, int OP>
__global__ __launch_bounds__(kBlock) void k_nlcg_vec(AtomArrays A, double beta, Vec<D> lfac,
                                                     ReduceBuf R) {
    using PT = typename PosStore<D>::T;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[kRedWidth];
#pragma unroll
    for (int k = 0; k < kRedWidth; k++) v[k] = 0.0;
    if (i < A.n) {
        const double m = A.m[i];
        const bool fixed = is_fixed(m);
        double f[D], a[D], vv[D];
#pragma unroll
        for (int k = 0; k < D; k++) {
            f[k] = A.f[k][i];
            a[k] = A.a[k][i];
            vv[k] = A.v[k][i];
        }
        bool wa = false, wv = false;
        if (OP == kNlcgAssignAV) {  // set_forces(true, ...): a = v = f
#pragma unroll
            for (int k = 0; k < D; k++) a[k] = vv[k] = fixed ? 0.0 : f[k];
            wa = wv = true;
        } else if (OP == kNlcgAssignA) {  // a = f
#pragma unroll
            for (int k = 0; k < D; k++) a[k] = fixed ? 0.0 : f[k];
            wa = true;
        } else if (OP == kNlcgAssignV) {  // v = a + v * beta
#pragma unroll
            for (int k = 0; k < D; k++) vv[k] = fixed ? 0.0 : a[k] + vv[k] * beta;
            wv = true;
        } else if (OP == kNlcgDescend) {  // v = a = f (fixed atoms untouched)
            if (!fixed) {
#pragma unroll
                for (int k = 0; k < D; k++) a[k] = vv[k] = f[k];
                wa = wv = true;
            }
        } else if (OP == kNlcgFixedA0) {  // a = 0 for fixed atoms
            if (fixed) {
#pragma unroll
                for (int k = 0; k < D; k++) a[k] = 0.0;
                wa = true;
            }
        }
        if (wa) {
#pragma unroll
            for (int k = 0; k < D; k++) A.a[k][i] = a[k];
        }
        if (wv) {
#pragma unroll
            for (int k = 0; k < D; k++) A.v[k][i] = vv[k];
        }
        if (!fixed) {
            v[0] = dotv<D>(f, f);
            v[1] = dotv<D>(f, a);
            v[2] = dotv<D>(f, vv);
            v[3] = dotv<D>(vv, vv);
            const Vec<D> x = PosStore<D>::load(static_cast<const PT *>(A.pos_in)[i]);
            double w[D];
#pragma unroll
            for (int k = 0; k < D; k++) w[k] = vv[k] + x.c[k] * lfac.c[k];
            v[4] = dotv<D>(w, w);
        }
    }
    block_partial<true, false, kBlock / kWave>(v, Top2{}, R);
}

hipError_t launch_nlcg_stepx(int ndim, const AtomArrays &A, const double *lfac, double dx,
                             hipStream_t s) {
    if (A.n <= 0) return hipSuccess;
    const dim3 g(reduce_blocks(A.n)), b(kBlock);
    if (ndim == 3) {
        Vec<3> l{{lfac[0], lfac[1], lfac[2]}};
        hipLaunchKernelGGL(k_nlcg_stepx<3>, g, b, 0, s, A, l, dx);
    } else {
        Vec<2> l{{lfac[0], lfac[1]}};
        hipLaunchKernelGGL(k_nlcg_stepx<2>, g, b, 0, s, A, l, dx);
    }
    return hipGetLastError();
}

template <int D>
static void nlcg_vec_d(int op, const AtomArrays &A, double beta, const Vec<D> &l,
                       const ReduceBuf &R, hipStream_t s) {
    const dim3 g(reduce_blocks(A.n)), b(kBlock);
    switch (op) {
        case kNlcgDots:
            hipLaunchKernelGGL((k_nlcg_vec<D, kNlcgDots>), g, b, 0, s, A, beta, l, R);
            break;
        case kNlcgAssignAV:
            hipLaunchKernelGGL((k_nlcg_vec<D, kNlcgAssignAV>), g, b, 0, s, A, beta, l, R);
            break;
        case kNlcgAssignA:
            hipLaunchKernelGGL((k_nlcg_vec<D, kNlcgAssignA>), g, b, 0, s, A, beta, l, R);
            break;
        case kNlcgAssignV:
            hipLaunchKernelGGL((k_nlcg_vec<D, kNlcgAssignV>), g, b, 0, s, A, beta, l, R);
            break;
        case kNlcgDescend:
            hipLaunchKernelGGL((k_nlcg_vec<D, kNlcgDescend>), g, b, 0, s, A, beta, l, R);
            break;
        default:
            hipLaunchKernelGGL((k_nlcg_vec<D, kNlcgFixedA0>), g, b, 0, s, A, beta, l, R);
            break;
    }
}

hipError_t launch_nlcg_vec(int ndim, int op, const AtomArrays &A, double beta,
                           const double *lfac, const ReduceBuf &R, hipStream_t s) {
    if (A.n <= 0) return hipSuccess;
    if (ndim == 3)
        nlcg_vec_d<3>(op, A, beta, Vec<3>{{lfac[0], lfac[1], lfac[2]}}, R, s);
    else
        nlcg_vec_d<2>(op, A, beta, Vec<2>{{lfac[0], lfac[1]}}, R, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_finalize_rt(true, false, reduce_blocks(A.n), R, s);
}

}  // namespace parm
