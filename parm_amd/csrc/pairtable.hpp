// pairtable.hpp -- host-side per-type-pair parameter tables, computed exactly
// as the reference's pair constructors compute them:
//   LJAttractRepulsePair(a1, a2)  src/interaction.hpp:1256-1269
//   RepulsionPair(a1, a2)         src/interaction.hpp:1532-1537
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"

namespace parm {

// Fills tab[t1 * ntypes + t2]; harm = every Repulsion exponent is 2.
inline void make_pair_table(int kind, int ntypes, const double *eps, const double *sigma,
                            const double *third, std::vector<PairParam> &tab,
                            bool &harm) {
    tab.assign((size_t)ntypes * ntypes, PairParam{});
    harm = true;
    for (int t1 = 0; t1 < ntypes; t1++) {
        for (int t2 = 0; t2 < ntypes; t2++) {
            PairParam p{};
            if (kind == kPairLJ) {
                // LJAttractRepulsePair(a1, a2)  (interaction.hpp:1256-1269)
                double e = eps[(size_t)t1 * ntypes + t2];
                double sig = (sigma[t1] + sigma[t2]) / 2.0;
                double cut = std::max(third[t1], third[t2]);
                double cut_energy;
                if (e <= 0) {
                    cut = 1;
                    cut_energy = 0;
                    e = std::fabs(e);
                } else {
                    double mid = (1 - std::pow(cut, -6));
                    cut_energy = e * (mid * mid);
                }
                p.eps = e;
                p.sig = sig;
                p.sig2 = sig * sig;
                p.cut2 = cut * cut;
                p.lim_lo = p.sig2 * p.cut2 * (1.0 - 1e-14);
                p.lim_hi = p.sig2 * p.cut2 * (1.0 + 1e-14);
                p.p5 = cut_energy;
                p.p6 = 12 * e;
                p.lim_f = (float)(p.sig2 * p.cut2);
            } else {
                // RepulsionPair(a1, a2)  (interaction.hpp:1532-1537)
                double e = std::sqrt(eps[t1] * eps[t2]);
                double sig = (sigma[t1] + sigma[t2]) / 2.0;
                double nexp = (third[t1] + third[t2]) / 2.0;
                p.eps = e;
                p.sig = sig;
                p.sig2 = sig * sig;
                p.cut2 = p.sig2;
                p.lim_lo = p.sig2;
                p.lim_hi = p.sig2;
                p.p5 = nexp;
                p.p6 = nexp - 1;
                p.lim_f = (float)p.sig2;
                if (nexp != 2.0) harm = false;
            }
            p.eps_f = (float)p.eps;
            p.sig2_f = (float)p.sig2;
            p.p5_f = (float)p.p5;
            p.p6_f = (float)p.p6;
            p.sig_f = (float)p.sig;
            p.isig_f = (float)(1.0 / p.sig);
            tab[(size_t)t1 * ntypes + t2] = p;
        }
    }
}

// List diameter of a type: A::max_size() (IEpsSigCutAtom: sigma * sigcut,
// src/interaction.hpp:1017; EpsSigExpAtom: sigma, src/interaction.hpp:1464).
inline double type_diameter(int kind, const double *sigma, const double *third, int t) {
    return kind == kPairLJ ? sigma[t] * third[t] : sigma[t];
}

}  // namespace parm
