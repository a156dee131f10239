// nlcg.hpp -- launchers of the device passes of the nonlinear
// conjugate-gradient minimiser (nlcg.hip).
#pragma once

#include "kernels.hpp"

namespace parm {

// The assignment a k_nlcg_vec pass applies before its sums.
enum NlcgOp : int {
    kNlcgDots = 0,      // none
    kNlcgAssignAV = 1,  // a = v = f (fixed atoms: 0)
    kNlcgAssignA = 2,   // a = f (fixed atoms: 0)
    kNlcgAssignV = 3,   // v = a + v * beta (fixed atoms: 0)
    kNlcgDescend = 4,   // a = v = f for atoms that are not fixed
    kNlcgFixedA0 = 5,   // a = 0 for fixed atoms
};

// Output fields of a k_nlcg_vec pass (ReduceBuf out[]): sums over atoms with a
// finite positive mass.
constexpr int kNlcgFF = 0, kNlcgFA = 1, kNlcgFV = 2, kNlcgVV = 3, kNlcgKin = 4;

// x <- x * lfac[k] + v * dx (per dimension k), every atom.
hipError_t launch_nlcg_stepx(int ndim, const AtomArrays &A, const double *lfac, double dx,
                             hipStream_t s);
// The assignment `op`, then the sums into R.out (lfac scales x in field 4).
hipError_t launch_nlcg_vec(int ndim, int op, const AtomArrays &A, double beta,
                           const double *lfac, const ReduceBuf &R, hipStream_t s);

}  // namespace parm
