// pairtable.hpp -- host-side per-type-pair parameter tables, computed exactly
// as the reference's pair constructors compute them.  Every pair form below
// runs on one of the two device kernels: the LJ kernel (force
// 12 eps s (s - 1) / r^2 with s = sigma^6 / r^6, energy eps (1 - s)^2 minus a
// shift, cut where r^2 / sigma^2 > cut^2) or the repulsion kernel
// (eps (1 - R/sigma)^n / n).  The forms differ only in how the pair's
// parameters are mixed from the two atoms':
//   LJAttractRepulsePair(IEpsSigCutAtom, ...)  src/interaction.hpp:1256-1269
//   RepulsionPair(EpsSigExpAtom, ...)          src/interaction.hpp:1532-1537
//   LennardJonesCutPair(EpsSigCutAtom, ...)    src/interaction.hpp:970-974, 249-253
//   LennardJonesCutPair(IEpsISigCutAtom, ...)  src/interaction.hpp:975-979
//   LJRepulsePair(EpsSigAtom, ...)             src/interaction.hpp:875-891, 119-147
//   RepulsionPair(IEpsISigExpAtom, ...)        src/interaction.hpp:1537-1542
#pragma once

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace parm {

// Pair forms (the reference's atom type + pair type combinations).  The
// first two share their numbers with the kernel kinds.
enum PairForm : int {
    kFormLJAttractRepulse = 1,  // NListed<IEpsSigCutAtom, LJAttractRepulsePair>
    kFormRepulsion = 2,         // NListed<EpsSigExpAtom, RepulsionPair>
    kFormLJCut = 3,             // NListed<EpsSigCutAtom, LennardJonesCutPair>
    kFormLJCutIndexed = 4,      // NListed<IEpsISigCutAtom, LennardJonesCutPair>
    kFormLJRepulse = 5,         // NListed<EpsSigAtom, LJRepulsePair>
    kFormRepulsionIndexed = 6,  // NListed<IEpsISigExpAtom, RepulsionPair>
};

inline int form_kernel_kind(int form) {
    return (form == kFormRepulsion || form == kFormRepulsionIndexed) ? kPairRepulsion
                                                                      : kPairLJ;
}

// Forms whose eps (and, for the indexed ones, sigma) arrive as ntypes x
// ntypes rows: the value for the pair (t1, t2) is rows[t1 * ntypes + t2].
inline bool form_eps_rows(int form) {
    return form == kFormLJAttractRepulse || form == kFormLJCutIndexed ||
           form == kFormRepulsionIndexed;
}
inline bool form_sigma_rows(int form) {
    return form == kFormLJCutIndexed || form == kFormRepulsionIndexed;
}

// The reference looks a row-indexed parameter up from the pair's first atom
// (the later-added one).  Each pair here is evaluated from both atoms, so a
// table must read the same both ways; "" if it does, else the reason.
inline std::string check_rows_symmetric(int form, int ntypes, const double *eps,
                                        const double *sigma) {
    for (int t1 = 0; t1 < ntypes; t1++)
        for (int t2 = 0; t2 < t1; t2++) {
            const size_t a = (size_t)t1 * ntypes + t2, b = (size_t)t2 * ntypes + t1;
            if (form_eps_rows(form) && !(eps[a] == eps[b]))
                return "epsilon rows are not symmetric (types " + std::to_string(t1) + ", " +
                       std::to_string(t2) + ")";
            if (form_sigma_rows(form) && !(sigma[a] == sigma[b]))
                return "sigma rows are not symmetric (types " + std::to_string(t1) + ", " +
                       std::to_string(t2) + ")";
        }
    return "";
}

// LJ-kernel parameters from the mixed eps, sigma, cut and the energy shift.
inline void lj_param(PairParam &p, double e, double sig, double cut, double shift) {
    p.eps = e;
    p.sig = sig;
    p.sig2 = sig * sig;
    p.cut2 = cut * cut;
    p.lim_lo = p.sig2 * p.cut2 * (1.0 - 1e-14);
    p.lim_hi = p.sig2 * p.cut2 * (1.0 + 1e-14);
    p.p5 = shift;
    p.p6 = 12 * e;
    p.lim_f = (float)(p.sig2 * p.cut2);
}

inline void rep_param(PairParam &p, double e, double sig, double nexp) {
    p.eps = e;
    p.sig = sig;
    p.sig2 = sig * sig;
    p.cut2 = p.sig2;
    p.lim_lo = p.sig2;
    p.lim_hi = p.sig2;
    p.p5 = nexp;
    p.p6 = nexp - 1;
    p.lim_f = (float)p.sig2;
}

// Fills tab[t1 * ntypes + t2] for one pair form.  eps / sigma are per type or
// ntypes x ntypes rows (form_eps_rows / form_sigma_rows); third is sigcut
// (LJ forms) or the exponent (repulsion forms), per type, unused by
// kFormLJRepulse.  harm = every repulsion exponent is 2.
inline void make_pair_table(int form, int ntypes, const double *eps, const double *sigma,
                            const double *third, std::vector<PairParam> &tab,
                            bool &harm) {
    tab.assign((size_t)ntypes * ntypes, PairParam{});
    harm = true;
    for (int t1 = 0; t1 < ntypes; t1++) {
        for (int t2 = 0; t2 < ntypes; t2++) {
            const size_t ix = (size_t)t1 * ntypes + t2;
            PairParam p{};
            switch (form) {
                case kFormLJAttractRepulse: {
                    // LJAttractRepulsePair(a1, a2)  (interaction.hpp:1256-1269)
                    double e = eps[ix];
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
                    lj_param(p, e, sig, cut, cut_energy);
                    break;
                }
                case kFormLJCut:
                case kFormLJCutIndexed: {
                    // LennardJonesCutPair: eps = sqrt(e1 e2), sig = (s1 + s2) / 2
                    // (EpsSigCutAtom) or the rows' entries (IEpsISigCutAtom);
                    // cut = max(sigcut).  LennardJonesCut (interaction.hpp:
                    // 249-253) shifts by eps ((1 - cut^-6)^2 - 1) from an
                    // energy eps ((1 - s)^2 - 1): the same as eps (1 - s)^2
                    // minus eps (1 - cut^-6)^2.
                    const bool ind = form == kFormLJCutIndexed;
                    double e = ind ? eps[ix] : std::sqrt(eps[t1] * eps[t2]);
                    double sig = ind ? sigma[ix] : (sigma[t1] + sigma[t2]) / 2;
                    double cut = std::max(third[t1], third[t2]);
                    double rsix = std::pow(cut, 6);
                    double mid = (1 - 1 / rsix);
                    lj_param(p, e, sig, cut, e * (mid * mid));
                    break;
                }
                case kFormLJRepulse: {
                    // LJRepulsePair: eps = sqrt(e1 e2), sig = (s1 + s2) / 2,
                    // cut at sigma (rsq > 1), no shift (LJRepulsive,
                    // interaction.hpp:126-147)
                    double e = std::sqrt(eps[t1] * eps[t2]);
                    double sig = (sigma[t1] + sigma[t2]) / 2;
                    lj_param(p, e, sig, 1.0, 0.0);
                    break;
                }
                case kFormRepulsion:
                case kFormRepulsionIndexed: {
                    // RepulsionPair(a1, a2)  (interaction.hpp:1532-1543)
                    const bool ind = form == kFormRepulsionIndexed;
                    double e = ind ? eps[ix] : std::sqrt(eps[t1] * eps[t2]);
                    double sig = ind ? sigma[ix] : (sigma[t1] + sigma[t2]) / 2.0;
                    double nexp = (third[t1] + third[t2]) / 2.0;
                    rep_param(p, e, sig, nexp);
                    if (nexp != 2.0) harm = false;
                    break;
                }
                default:
                    break;
            }
            p.eps_f = (float)p.eps;
            p.sig2_f = (float)p.sig2;
            p.p5_f = (float)p.p5;
            p.p6_f = (float)p.p6;
            p.sig_f = (float)p.sig;
            p.isig_f = (float)(1.0 / p.sig);
            tab[ix] = p;
        }
    }
}

// List diameter of a type: A::max_size() (IEpsSigCutAtom: sigma * sigcut,
// src/interaction.hpp:1017; EpsSigExpAtom: sigma, src/interaction.hpp:1464).
// Used by the batched path, whose systems add every atom with these forms.
inline double type_diameter(int form, const double *sigma, const double *third, int t) {
    return form == kFormLJAttractRepulse ? sigma[t] * third[t] : sigma[t];
}

}  // namespace parm
