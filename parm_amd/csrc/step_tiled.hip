// step_tiled.hip -- pair forces over the tiled list (tiles.hpp), fused with
// the integrator stages exactly as k_force in step.hip.
//
// Reference behaviour restated (paths relative to the reference root):
//   NListed<A,P>::set_forces / energy / pressure / stress
//                                   src/interaction.hpp:2153-2291
//   LJAttractRepulsePair::forces/energy  src/interaction.hpp:1270-1298
//   RepulsionPair::forces/energy         src/interaction.hpp:1538-1556
//
// One workgroup per tile, TPA lanes per atom (kWave * TPA threads).  The
// tile's stencil (its runs, precomputed at the rebuild) is staged in LDS once
// per launch; each list entry is a staged slot, so a neighbour position is an
// LDS read instead of a gather from memory, and the list itself is 2 bytes
// per entry.  Three staging modes:
//   RAW   (FP64 repulsion): the positions themselves; each pair takes the
//         reference's minimum image diff(x_i, x_j), so every pair force is
//         the untiled kernel's bit for bit.
//   REL64 (FP64 LJ, mixed repulsion): the FP64 minimum-image offset of each
//         atom from the tile's first atom.  A pair displacement is the
//         difference of two offsets (both in [-L/2, L/2], so one fold gives
//         the minimum image; when every offset of the tile is under L/4 the
//         difference already is, and the fold is skipped).  It differs from
//         diff(x_i, x_j) only by rounding (~1e-16 relative); the LJ in/out
//         decision within 1e-14 of the cutoff is taken on the exact d2, so the
//         set of interacting pairs is the reference's.
//   REL32 (mixed LJ): the same offsets rounded to FP32 and differenced in
//         FP32 (half the LDS traffic; the rounding of offsets of a few sigma
//         costs ~1e-6 sigma in r).  Repulsion does not use it: its force
//         eps (1 - R/sigma)/sigma cancels at small overlaps.
// Mixed precision decides in/out within 1e-5 of the cutoff in FP64 on the
// exact d2, as the untiled mixed path does.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "pairs.hpp"
#include "stepcore.hpp"
#include "tiles.hpp"

namespace parm {

// FP64 LJAttractRepulsePair on a displacement r that may differ from the
// reference's by rounding; exact_d2() is the reference's d2, used only within
// 1e-14 of the cutoff (where lim_lo / lim_hi bracket it).
template <int D, bool ENERGY, typename Exact>
__device__ __forceinline__ void lj64x(const PairParam &pp, const Vec<D> &r,
                                      const Exact &exact_d2, double (&f)[D], PairAcc &acc) {
    // fused d2: within an ulp of the reference's (x*x + y*y) + z*z, far
    // inside the band in which exact_d2 decides
    double d2 = r.c[0] * r.c[0];
#pragma unroll
    for (int k = 1; k < D; k++) d2 = fma(r.c[k], r.c[k], d2);
    if (!(d2 < pp.lim_lo)) {
        if (d2 > pp.lim_hi) return;
        if (exact_d2() / pp.sig2 > pp.cut2) return;
    }
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

// x folded into the primary image, x - L rint(x / L): one rounding, of
// magnitude ulp(L / 2), whatever image the unwrapped x is in.
template <int D>
__device__ __forceinline__ Vec<D> fold_image(const Vec<D> &x, const Box &b) {
    Vec<D> r;
#pragma unroll
    for (int k = 0; k < D; k++) r.c[k] = fma(-b.L[k], rint(x.c[k] * b.invL[k]), x.c[k]);
    return r;
}

// Minimum-image offset of x from the origin o (o already folded):
// d - L rint(d / L) with d = x - o.  `far` is set when x lies more than a box
// length from o along some dimension: x - o then rounds at its own
// magnitude, and offset_folded (x folded into the primary image first, every
// rounding at ulp(L)) should be taken instead.  The staged offsets need not
// round like the reference's remainder (they are differenced again).
template <int D>
__device__ __forceinline__ Vec<D> offset_from(const Vec<D> &x, const Vec<D> &o, const Box &b,
                                              bool &far) {
    Vec<D> r;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const double d = x.c[k] - o.c[k];
        far |= fabs(d) > b.L[k];
        r.c[k] = fma(-b.L[k], rint(d * b.invL[k]), d);
    }
    return r;
}

template <int D>
__device__ __forceinline__ Vec<D> offset_folded(const Vec<D> &x, const Vec<D> &o,
                                                const Box &b) {
    const Vec<D> xf = fold_image<D>(x, b);
    Vec<D> r;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const double d = xf.c[k] - o.c[k];
        r.c[k] = fma(-b.L[k], rint(d * b.invL[k]), d);
    }
    return r;
}

// List words per lane read ahead of the pair loop (two entries a word), and
// the entries the one-type LJ loop evaluates side by side.
constexpr int kChunkWords = 4;
constexpr int kGroup = 4;
static_assert(kChunkWords % (kGroup / 2) == 0, "groups tile a chunk");

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE, int TPA>
__global__ __launch_bounds__(kWave *TPA) void k_force_tiled(AtomArrays A, ListView Lv,
                                                            PairView P, Box box, StepCoef c,
                                                            ReduceBuf R) {
    using PT = typename PosStore<D>::T;
    constexpr bool ENERGY = STAGE == kStageSums;
    constexpr bool RAW = PREC == kFP64 && KIND != kPairLJ;
    constexpr bool REL32 = PREC != kFP64 && KIND == kPairLJ;
    // COMPACT (one pair type, offsets staged): a slot is the D offsets only
    // (FP64: 8 D bytes, FP32: 4 D bytes), no type, no atom index (the rare
    // exact-d2 path finds the atom from the runs): less LDS per workgroup,
    // more workgroups per CU.
    constexpr bool COMPACT = SINGLE && !RAW;
    constexpr int NT = kWave * TPA;
    extern __shared__ double4 st64[];
    __shared__ int2 runs[kRunSlots];  // run_entry form (tiles.hpp)
    __shared__ int4 hdr;
    float4 *st32 = reinterpret_cast<float4 *>(st64);
    double *stc64 = reinterpret_cast<double *>(st64);
    float *stc32 = reinterpret_cast<float *>(st64);
    int32_t *stj = REL32 ? reinterpret_cast<int32_t *>(st32 + Lv.smax)
                         : reinterpret_cast<int32_t *>(st64 + Lv.smax);

    const int t = xcd_block(blockIdx.x, gridDim.x);
    if (threadIdx.x < kRunSlots) runs[threadIdx.x] = run_entry(Lv.truns, t, threadIdx.x);
    if (threadIdx.x == 0) hdr = Lv.tiles[t];
    __syncthreads();
    const int a0 = hdr.x, tcnt = hdr.y;
    const int S = hdr.z >= 0 ? hdr.w : 0;
    const PT *pos = static_cast<const PT *>(A.pos_in);
    // the first wavefront updates the tile's atoms at the end (one lane per
    // atom); their inputs are loaded now, to arrive during the pair loop
    AtomIn<D> ain;
    const bool upd = STAGE != kStageSums && threadIdx.x < (unsigned)tcnt;
    if (upd) atom_load<D, STAGE>(A, a0 + (int)threadIdx.x, ain);
    // this thread's atom and lane, and what its pair loop reads from memory
    // (position, membership, entry count, the first chunk of list words),
    // also loaded now: their latency hides behind the staging.  The first
    // chunk is read whatever the lane's count (within the list array's
    // wcap words per lane); words past the count are never used.
    const int li = threadIdx.x / TPA, sub = threadIdx.x % TPA;
    const int i = a0 + li;
    const bool act = li < tcnt;
    Vec<D> xi;
#pragma unroll
    for (int k = 0; k < D; k++) xi.c[k] = 0.0;
    int rk = -1, nlane = 0;
    const uint32_t *wp = Lv.nbr + ((size_t)i * TPA + sub);
    const size_t wstride = (size_t)Lv.npad * TPA;
    uint32_t cur[kChunkWords];
#pragma unroll
    for (int q = 0; q < kChunkWords; q++) cur[q] = 0u;
    if (act) {
        xi = PosStore<D>::load(pos[i]);
        if (hdr.z >= 0) {
            rk = A.rank[i];
            nlane = Lv.lcount[(size_t)i * TPA + sub];
#pragma unroll
            for (int q = 0; q < kChunkWords; q++)
                if (q < Lv.wcap) cur[q] = wp[(size_t)q * wstride];
        }
    }
    // the tile origin: loaded now, folded once the stencil's loads are issued
    PT xo_raw{};
    if (!RAW && S > 0) xo_raw = pos[a0];
    Vec<D> xo;
#pragma unroll
    for (int k = 0; k < D; k++) xo.c[k] = 0.0;
    bool xo_ready = false;
    int wide = 0;  // some offset reaches L/4: pair differences need the fold
    // kStageBatch slots per thread at a time: their atom indices first (each
    // an independent binary lifting over the runs), then all their position
    // (and type) loads in flight together, then the stores
    constexpr int kStageBatch = 5;
    for (int s0 = threadIdx.x; s0 < S; s0 += kStageBatch * NT) {
        int jb[kStageBatch];
        PT pb[kStageBatch];
        int tb[kStageBatch];
        {
            int sb[kStageBatch];
#pragma unroll
            for (int q = 0; q < kStageBatch; q++) sb[q] = s0 + q * NT;
            slot_atoms<kStageBatch>(runs, sb, jb);
        }
#pragma unroll
        for (int q = 0; q < kStageBatch; q++) {
            if (s0 + q * NT < S) {
                pb[q] = pos[jb[q]];
                tb[q] = SINGLE ? 0 : A.type[jb[q]];
            }
        }
        if (!RAW && !xo_ready) {
            xo = fold_image<D>(PosStore<D>::load(xo_raw), box);
            xo_ready = true;
        }
#pragma unroll
        for (int q = 0; q < kStageBatch; q++) {
        const int s = s0 + q * NT;
        if (s >= S) break;
        const int j = jb[q];
        const Vec<D> xj = PosStore<D>::load(pb[q]);
        const int tj = tb[q];
        Vec<D> v = xj;
        if (!RAW) {
            bool far = false;
            v = offset_from<D>(xj, xo, box, far);
            // atoms several box lengths out are rare: one wave-uniform test
            if (__ballot(far) != 0ull) {
                if (far) v = offset_folded<D>(xj, xo, box);
            }
#pragma unroll
            for (int k = 0; k < D; k++) wide |= fabs(v.c[k]) >= 0.5 * box.halfL[k];
            if (!COMPACT) stj[s] = j;
        }
        if (COMPACT) {
#pragma unroll
            for (int k = 0; k < D; k++) {
                if (REL32)
                    stc32[s * D + k] = (float)v.c[k];
                else
                    stc64[s * D + k] = v.c[k];
            }
        } else if (REL32) {
            float4 q;
            q.x = (float)v.c[0];
            q.y = (float)v.c[1];
            q.z = D == 3 ? (float)v.c[2] : 0.0f;
            q.w = __int_as_float(tj);
            st32[s] = q;
        } else {
            double4 q;
            q.x = v.c[0];
            q.y = v.c[1];
            if (D == 3) {
                q.z = v.c[2];
                q.w = (double)tj;
            } else {
                q.z = (double)tj;
                q.w = 0.0;
            }
            st64[s] = q;
        }
        }
    }
    // (a thread with no slot to stage still needs the origin)
    if (!RAW && !xo_ready) xo = fold_image<D>(PosStore<D>::load(xo_raw), box);
    const bool fold = __syncthreads_or(wide) != 0;

    PairAcc acc;
    double fi[D];
    float f32[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        fi[k] = 0.0;
        f32[k] = 0.0f;
    }
    const bool member = act && hdr.z >= 0 && rk >= 0;
    if (member) {
        const PairParam *prow = P.tab;
        PairParam p0;
        if (SINGLE)
            p0 = P.p0;
        else
            prow = P.tab + (size_t)A.type[i] * P.ntypes;
        Vec<D> ri64 = xi;
        float ri[D], Lf[D], hLf[D];
#pragma unroll
        for (int k = 0; k < D; k++) ri[k] = Lf[k] = hLf[k] = 0.0f;
        if (!RAW) {
            ri64 = offset_folded<D>(xi, xo, box);
#pragma unroll
            for (int k = 0; k < D; k++) {
                ri[k] = (float)ri64.c[k];
                Lf[k] = (float)box.L[k];
                hLf[k] = (float)box.halfL[k];
            }
        }
        // the reference's d2 (for in/out decisions at the cutoff)
        auto exact_d2 = [&](int s) {
            int j;
            if constexpr (COMPACT) {
                j = slot_atom(runs, s);
            } else {
                j = stj[s];
            }
            const Vec<D> r = diff<D>(xi, PosStore<D>::load(pos[j]), box);
            return sqnorm<D>(r);
        };
        auto pair = [&](int s, auto fold_c) {
            constexpr bool FOLD = decltype(fold_c)::value;
            if constexpr (RAW) {
                const double4 q = st64[s];
                Vec<D> xj;
                xj.c[0] = q.x;
                xj.c[1] = q.y;
                if constexpr (D == 3) xj.c[2] = q.z;
                const int tj = SINGLE ? 0 : (int)(D == 3 ? q.w : q.z);
                const Vec<D> r = diff<D>(xi, xj, box);
                visit<D, KIND, PREC, ENERGY, HARM>(SINGLE ? p0 : prow[tj], r, fi, f32, acc);
            } else if constexpr (REL32) {
                float qc[3] = {0.0f, 0.0f, 0.0f};
                int tj = 0;
                if constexpr (COMPACT) {
#pragma unroll
                    for (int k = 0; k < D; k++) qc[k] = stc32[s * D + k];
                } else {
                    const float4 q = st32[s];
                    qc[0] = q.x;
                    qc[1] = q.y;
                    qc[2] = q.z;
                    tj = __float_as_int(q.w);
                }
                float rf[D];
#pragma unroll
                for (int k = 0; k < D; k++) {
                    float d = ri[k] - qc[k];
                    if (FOLD) d = d > hLf[k] ? d - Lf[k] : (d < -hLf[k] ? d + Lf[k] : d);
                    rf[k] = d;
                }
                lj32f<D, ENERGY>(SINGLE ? p0 : prow[tj], rf, [&]() { return exact_d2(s); },
                                 f32, acc);
            } else {
                double qc[3] = {0.0, 0.0, 0.0};
                int tj = 0;
                if constexpr (COMPACT) {
#pragma unroll
                    for (int k = 0; k < D; k++) qc[k] = stc64[s * D + k];
                } else {
                    const double4 q = st64[s];
                    qc[0] = q.x;
                    qc[1] = q.y;
                    qc[2] = q.z;
                    tj = (int)(D == 3 ? q.w : q.z);
                }
                Vec<D> r;
#pragma unroll
                for (int k = 0; k < D; k++) {
                    double d = ri64.c[k] - qc[k];
                    if (FOLD)
                        d = d > box.halfL[k] ? d - box.L[k]
                                             : (d < -box.halfL[k] ? d + box.L[k] : d);
                    r.c[k] = d;
                }
                const PairParam &pp = SINGLE ? p0 : prow[tj];
                if constexpr (PREC == kFP64) {
                    lj64x<D, ENERGY>(pp, r, [&]() { return exact_d2(s); }, fi, acc);
                } else {
                    float rf[D];
#pragma unroll
                    for (int k = 0; k < D; k++) rf[k] = (float)r.c[k];
                    rep32f<D, ENERGY, HARM>(pp, rf, [&]() { return exact_d2(s); }, f32, acc);
                }
            }
        };
        // kGroup entries (kGroup / 2 words) of the one-type LJ list at once,
        // branch-free but for the in/out decision within a hair of the
        // cutoff: their staged offsets read together, their force factors
        // evaluated side by side, then added in list order (an entry outside
        // the cutoff, or past the lane's count, adds 0 * r, which leaves the
        // sum as it is; a slot past the count reads slot 0, which holds a
        // staged offset whenever any entry exists).
        auto group = [&](const uint32_t *wd, int e, auto fold_c) {
            constexpr bool FOLD = decltype(fold_c)::value;
            int sl[kGroup];
            bool ok[kGroup], in[kGroup], band[kGroup];
#pragma unroll
            for (int g = 0; g < kGroup; g++) {
                const uint32_t w = wd[g >> 1];
                ok[g] = e + g < nlane;
                sl[g] = ok[g] ? (int)((g & 1) ? (w >> 16) : (w & 0xffffu)) : 0;
            }
            if constexpr (PREC == kFP64) {
                double r[kGroup][D], fs[kGroup];
#pragma unroll
                for (int g = 0; g < kGroup; g++) {
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        double d = ri64.c[k] - stc64[sl[g] * D + k];
                        if (FOLD)
                            d = d > box.halfL[k] ? d - box.L[k]
                                                 : (d < -box.halfL[k] ? d + box.L[k] : d);
                        r[g][k] = d;
                    }
                    double d2 = r[g][0] * r[g][0];
#pragma unroll
                    for (int k = 1; k < D; k++) d2 = fma(r[g][k], r[g][k], d2);
                    const bool lo = d2 < p0.lim_lo;
                    band[g] = ok[g] && !lo && !(d2 > p0.lim_hi);
                    in[g] = ok[g] && lo;
                    double rsix, fmagR;
                    fs[g] = lj64_fs(p0, d2, rsix, fmagR);
                }
                bool anyband = false;
#pragma unroll
                for (int g = 0; g < kGroup; g++) anyband |= band[g];
                if (anyband) {
#pragma unroll
                    for (int g = 0; g < kGroup; g++)
                        if (band[g]) in[g] = !(exact_d2(sl[g]) / p0.sig2 > p0.cut2);
                }
#pragma unroll
                for (int g = 0; g < kGroup; g++) {
                    const double m = in[g] ? fs[g] : 0.0;
#pragma unroll
                    for (int k = 0; k < D; k++) fi[k] = fma(m, r[g][k], fi[k]);
                }
            } else {
                float r[kGroup][D], fs[kGroup];
#pragma unroll
                for (int g = 0; g < kGroup; g++) {
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        float d = ri[k] - stc32[sl[g] * D + k];
                        if (FOLD) d = d > hLf[k] ? d - Lf[k] : (d < -hLf[k] ? d + Lf[k] : d);
                        r[g][k] = d;
                    }
                    float d2 = r[g][0] * r[g][0];
#pragma unroll
                    for (int k = 1; k < D; k++) d2 = fmaf(r[g][k], r[g][k], d2);
                    const bool lo = !(d2 > p0.lim_f * 0.99999f);
                    band[g] = ok[g] && !lo && !(d2 > p0.lim_f * 1.00001f);
                    in[g] = ok[g] && lo;
                    float rsix, fmagR;
                    fs[g] = lj32_fs(p0, d2, rsix, fmagR);
                }
                bool anyband = false;
#pragma unroll
                for (int g = 0; g < kGroup; g++) anyband |= band[g];
                if (anyband) {
#pragma unroll
                    for (int g = 0; g < kGroup; g++) {
                        if (band[g]) {
                            const double d2d = exact_d2(sl[g]);
                            in[g] = d2d < p0.lim_lo ||
                                    (!(d2d > p0.lim_hi) && !(d2d / p0.sig2 > p0.cut2));
                        }
                    }
                }
#pragma unroll
                for (int g = 0; g < kGroup; g++) {
                    const float m = in[g] ? fs[g] : 0.0f;
#pragma unroll
                    for (int k = 0; k < D; k++) f32[k] = fmaf(m, r[g][k], f32[k]);
                }
            }
        };
        // two entries per 32-bit word; for the one-type LJ forms the lane's
        // words stream from memory kChunkWords at a time, each chunk loaded
        // while the one before it is evaluated (the first was loaded before
        // the staging)
        auto loop = [&](auto fold_c) {
            if constexpr (COMPACT && !ENERGY && KIND == kPairLJ) {
                const int nw = (nlane + 1) >> 1;
                for (int w0 = 0; w0 < nw; w0 += kChunkWords) {
                    uint32_t nxt[kChunkWords];
#pragma unroll
                    for (int q = 0; q < kChunkWords; q++) {
                        const int w = w0 + kChunkWords + q;
                        nxt[q] = w < nw ? wp[(size_t)w * wstride] : 0u;
                    }
#pragma unroll
                    for (int q = 0; q < kChunkWords; q += kGroup / 2)
                        if (2 * (w0 + q) < nlane) group(cur + q, 2 * (w0 + q), fold_c);
#pragma unroll
                    for (int q = 0; q < kChunkWords; q++) cur[q] = nxt[q];
                }
            } else {
                // the other forms one entry at a time, the word two ahead
                // loaded while the current one is evaluated
                uint32_t w0 = cur[0], w1 = cur[1];
                for (int m = 0; m < nlane; m += 2) {
                    const uint32_t c = w0;
                    w0 = w1;
                    if (m + 4 < nlane) w1 = wp[(size_t)((m >> 1) + 2) * wstride];
                    pair((int)(c & 0xffffu), fold_c);
                    if (m + 1 < nlane) pair((int)(c >> 16), fold_c);
                }
            }
        };
        if (fold)
            loop(std::integral_constant<bool, true>());
        else
            loop(std::integral_constant<bool, false>());
    }
    if (PREC != kFP64) {
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] += (double)f32[k];
    }
    // the atom's TPA lanes summed (xor butterfly; within a quad by DPP)
    if constexpr (TPA == 2) {
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] += dpp_f64<kDppQuadXor1>(fi[k]);
    } else if constexpr (TPA >= 4) {
#pragma unroll
        for (int k = 0; k < D; k++) fi[k] = quad_sum(fi[k]);
#pragma unroll
        for (int o = 4; o < TPA; o <<= 1) {
#pragma unroll
            for (int k = 0; k < D; k++) fi[k] += __shfl_xor(fi[k], o);
        }
    }
    // The integrator's per-atom memory traffic is done by the first
    // wavefront, one lane per atom (64 consecutive atoms: coalesced), with
    // the summed forces passed through LDS.
    Top2 top;
    double nonfinite = 0.0;
    if (STAGE != kStageSums) {
        __shared__ double fsum[D][kTileAtoms];
        if (act && sub == 0) {
#pragma unroll
            for (int k = 0; k < D; k++) fsum[k][li] = fi[k];
        }
        __syncthreads();
        if (upd) {
            double fa[D];
#pragma unroll
            for (int k = 0; k < D; k++) fa[k] = fsum[k][threadIdx.x];
            atom_store<D, STAGE>(A, a0 + (int)threadIdx.x, fa, ain, c, top, nonfinite);
        }
    }
    group_partials<D, STAGE, ENERGY, TPA, true>(acc, top, nonfinite, R);
}

// ======================================================== host launchers
size_t tiled_force_lds(int ndim, int prec, int kind, bool single, int smax) {
    const bool raw = prec == kFP64 && kind == kPairRepulsion;
    const bool rel32 = prec != kFP64 && kind == kPairLJ;
    if (raw) return (size_t)smax * sizeof(double4);
    if (single && kind != kPairNone)
        return (size_t)smax * ndim * (rel32 ? sizeof(float) : sizeof(double));
    if (rel32) return (size_t)smax * (sizeof(float4) + sizeof(int32_t));
    return (size_t)smax * (sizeof(double4) + sizeof(int32_t));
}

namespace {

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE, int TPA>
void tiled_tt(const AtomArrays &A, const ListView &Lv, const PairView &P, const Box &box,
              const StepCoef &c, const ReduceBuf &R, hipStream_t s) {
    constexpr bool DRIFT =
        (STAGE == kStageVerletKickDrift || STAGE == kStageDampedKickDrift);
    constexpr bool ENERGY = STAGE == kStageSums;
    const size_t lds = tiled_force_lds(D, PREC, KIND, SINGLE, Lv.smax);
    hipLaunchKernelGGL((k_force_tiled<D, KIND, PREC, SINGLE, HARM, STAGE, TPA>),
                       dim3(Lv.ntiles), dim3(kWave * TPA), lds, s, A, Lv, P, box, c, R);
    if (R.done) (void)hipEventRecord(R.done, s);
    if (ENERGY || DRIFT) (void)launch_finalize_rt(ENERGY, DRIFT, Lv.ntiles, R, s);
}

template <int D, int KIND, int PREC, bool SINGLE, bool HARM, int STAGE>
void tiled_t(const AtomArrays &A, const ListView &Lv, const PairView &P, const Box &box,
             const StepCoef &c, const ReduceBuf &R, hipStream_t s) {
    switch (Lv.tpa) {
    case 1:
        tiled_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 1>(A, Lv, P, box, c, R, s);
        break;
    case 2:
        tiled_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 2>(A, Lv, P, box, c, R, s);
        break;
    case 8:
        tiled_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 8>(A, Lv, P, box, c, R, s);
        break;
    default:
        tiled_tt<D, KIND, PREC, SINGLE, HARM, STAGE, 4>(A, Lv, P, box, c, R, s);
        break;
    }
}

template <int D, int KIND, int PREC, bool SINGLE, bool HARM>
void tiled_s(int stage, const AtomArrays &A, const ListView &Lv, const PairView &P,
             const Box &box, const StepCoef &c, const ReduceBuf &R, hipStream_t s) {
    switch (stage) {
#define PARM_TSTAGE(S)                                                    \
    case S:                                                               \
        tiled_t<D, KIND, PREC, SINGLE, HARM, S>(A, Lv, P, box, c, R, s); \
        break;
        PARM_TSTAGE(kStageForceOnly)
        PARM_TSTAGE(kStageForceA)
        PARM_TSTAGE(kStageVerletKick)
        PARM_TSTAGE(kStageVerletKickDrift)
        PARM_TSTAGE(kStageDampedKick)
        PARM_TSTAGE(kStageDampedKickDrift)
        PARM_TSTAGE(kStageSums)
        PARM_TSTAGE(kStageForceAdd)
#undef PARM_TSTAGE
    }
}

template <int D, int KIND, int PREC>
void tiled_k(int stage, const AtomArrays &A, const ListView &Lv, const PairView &P,
             const Box &box, const StepCoef &c, const ReduceBuf &R, hipStream_t s) {
    if constexpr (KIND == kPairRepulsion) {
        if (P.harmonic) {
            if (P.single)
                tiled_s<D, KIND, PREC, true, true>(stage, A, Lv, P, box, c, R, s);
            else
                tiled_s<D, KIND, PREC, false, true>(stage, A, Lv, P, box, c, R, s);
            return;
        }
    }
    if (P.single)
        tiled_s<D, KIND, PREC, true, false>(stage, A, Lv, P, box, c, R, s);
    else
        tiled_s<D, KIND, PREC, false, false>(stage, A, Lv, P, box, c, R, s);
}

template <int D>
void tiled_d(int precision, int stage, const AtomArrays &A, const ListView &Lv,
             const PairView &P, const Box &box, const StepCoef &c, const ReduceBuf &R,
             hipStream_t s) {
    if (P.kind == kPairRepulsion) {
        if (precision == kFP64)
            tiled_k<D, kPairRepulsion, kFP64>(stage, A, Lv, P, box, c, R, s);
        else
            tiled_k<D, kPairRepulsion, kMixed>(stage, A, Lv, P, box, c, R, s);
    } else {
        if (precision == kFP64)
            tiled_k<D, kPairLJ, kFP64>(stage, A, Lv, P, box, c, R, s);
        else
            tiled_k<D, kPairLJ, kMixed>(stage, A, Lv, P, box, c, R, s);
    }
}

}  // namespace

hipError_t launch_force_tiled(int ndim, int precision, int stage, bool energy,
                              const AtomArrays &A, const ListView &Lv, const PairView &P,
                              const Box &box, const StepCoef &c, const ReduceBuf &R,
                              hipStream_t s) {
    if (energy != (stage == kStageSums)) return hipErrorInvalidValue;
    if (A.n <= 0 || Lv.ntiles <= 0) return hipSuccess;
    if (ndim == 3)
        tiled_d<3>(precision, stage, A, Lv, P, box, c, R, s);
    else
        tiled_d<2>(precision, stage, A, Lv, P, box, c, R, s);
    return hipGetLastError();
}

}  // namespace parm
