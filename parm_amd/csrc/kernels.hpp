// kernels.hpp -- host-side launchers of the device kernels (internal API of
// libparm_hip.so; the public C ABI is include/parm_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include <cstdint>

#include "common.hpp"

namespace parm {

// Periodic cell grid used by the list rebuild.
struct CellGrid {
    int nc[3];       // cells per dimension (1 when the box is under 3 list radii)
    int ncell;       // product of nc over the active dimensions
    double invL[3];  // 1 / box edge
};

// ---------------------------------------------------------------- nlist.hip
int radix_blocks(int n);
hipError_t launch_cell_keys(int ndim, int n, const void *pos, const int32_t *rank,
                            const CellGrid &g, uint32_t *keys, int32_t *vals,
                            hipStream_t s);
hipError_t radix_sort(int n, int bits, uint32_t *&keys, int32_t *&vals,
                      uint32_t *&tmp_keys, int32_t *&tmp_vals, uint32_t *hist,
                      hipStream_t s);
template <typename T>
hipError_t gather(int n, const int32_t *perm, const T *src, T *dst, hipStream_t s);
// One permutation launch over every per-slot column: slot i of each
// destination takes slot perm[i] of its source.
constexpr int kPermD = 11, kPermI = 3;
struct PermuteSet {
    const void *pos_src[2];
    void *pos_dst[2];
    const double *d_src[kPermD];
    double *d_dst[kPermD];
    const int32_t *i_src[kPermI];
    int32_t *i_dst[kPermI];
    int nd, ni;
};
hipError_t launch_permute(int ndim, int n, const int32_t *perm, const PermuteSet &P,
                          hipStream_t s);
hipError_t launch_cell_bounds(int n, int ncell, const uint32_t *keys, int32_t *start,
                              hipStream_t s);
// Outcome of one list build, stored by the build's last launch in pinned
// host memory: the sum and the largest of the per-atom entry counts, the
// flag word of the build (then cleared on the device), and for a
// speculative tiled build the tile plan it ran against.
struct BuildReport {
    unsigned long long sum, max;
    int ntiles, smax, bad;
    uint32_t flag;
};

// NeighborList::ignore exclusions (src/trackers.hpp:190-193): excluded
// partners of each atom as a CSR over original atom indices (both directions
// stored, partners sorted); orig maps a device slot to its original index.
// any == 0 disables the test.
struct ExclView {
    const int32_t *orig;
    const int32_t *off;   // natoms + 1
    const int32_t *part;  // off[natoms] partners
    int any;
};

// Builds the full list (and xref = pos); stats[0] = sum of counts, stats[1] =
// max count; part holds 2 * build_blocks(n) scratch words.  report: a
// device-visible pinned BuildReport, or null.
int build_blocks(int n);
hipError_t launch_build_list(int ndim, int n, const void *pos, void *xw,
                             const int32_t *rank, const double *diam, const int32_t *type,
                             const uint32_t *keys, const int32_t *cstart,
                             const CellGrid &g, const Box &box, double skin, double dmax,
                             int type_bits, int tpa, int cap, int npad, uint32_t *nbr,
                             int32_t *ncount, uint16_t *lcount, void *xref,
                             unsigned long long *part, unsigned long long *stats,
                             uint32_t *flag, const ExclView &ex, BuildReport *report,
                             hipStream_t s);

// Tiled layout (tiles.hpp).  launch_tiles fills tiles (cap_tiles records)
// and info[0] = tiles, info[1] = largest staged stencil, info[2] != 0 when the
// layout does not apply.
hipError_t launch_tiles(int ndim, int n, const CellGrid &g, const uint32_t *keys,
                        const int32_t *cstart, int cap_tiles, int4 *tiles, int2 *truns,
                        int *info, hipStream_t s);
size_t tiled_build_lds(int smax);
// spec == null: ntiles and smax are the plan's.  spec == the plan's info
// (launch_tiles, not yet read by the host): ntiles and smax are bounds;
// tiles past info[0] do nothing, and a tile with more than smax staged slots
// (or none that fit) sets kFlagReplan instead of building.  uni: every list
// member has the same diameter.
hipError_t launch_build_tiled(int ndim, int n, int ntiles, int smax, const int4 *tiles,
                              const int2 *truns, const void *pos, void *xw, const int32_t *rank,
                              const double *diam, const int32_t *type, int type_bits,
                              const uint32_t *keys, const int32_t *cstart, const CellGrid &g,
                              const Box &box, double skin, double dmax, int tpa, int cap,
                              int npad, bool out32, void *nbr, int32_t *ncount,
                              uint16_t *lcount, void *xref, unsigned long long *part,
                              unsigned long long *stats, uint32_t *flag, const int *spec,
                              int uni, const ExclView &ex, BuildReport *report,
                              hipStream_t s);
hipError_t launch_tiled_to_global(int ndim, int ntiles, const int4 *tiles,
                                  const uint32_t *keys, const int32_t *cstart,
                                  const CellGrid &g, const uint16_t *nb16,
                                  const uint16_t *lcount, const int32_t *type, int type_bits,
                                  int tpa, int npad, uint32_t *out, hipStream_t s);

// -------------------------------------------------------------- step.hip
enum PairKind : int { kPairNone = 0, kPairLJ = 1, kPairRepulsion = 2 };
enum Precision : int { kFP64 = 0, kMixed = 1 };

// Per type-pair parameters, precomputed on the host exactly as the
// reference's pair constructors compute them (src/interaction.hpp:1256-1269,
// 1532-1537).  Indexed [t1 * ntypes + t2].
struct PairParam {
    double eps;        // LJ: |eps| ; Repulsion: sqrt(e1 e2)
    double sig;        // (s1 + s2) / 2
    double sig2;       // sig * sig
    double cut2;       // LJ: cut * cut (in sigma units) ; Repulsion: sig * sig
    double lim_lo;     // d2 below: surely inside the cutoff
    double lim_hi;     // d2 above: surely outside
    double p5;         // LJ: cut_energy      ; Repulsion: exponent n
    double p6;         // LJ: 12 * eps        ; Repulsion: n - 1
    float eps_f, sig2_f, p5_f, p6_f;  // single-precision copies (mixed mode)
    float lim_f, sig_f, isig_f, pad_f;
};

// Per-system views handed to the step kernels.
struct AtomArrays {
    int n;                 // slots
    const void *pos_in;    // PosStore<D>::T*
    void *pos_out;         // fused steps write the drifted positions here
    double *v[3], *a[3], *f[3];
    const double *m;
    const int32_t *type;   // pair-table type of each slot
    const int32_t *rank;   // < 0: not a list member (no pair forces, no skin test)
    const void *xref;      // positions at the last rebuild
};

struct ListView {
    const uint32_t *nbr;
    const int32_t *ncount;
    const uint16_t *lcount;  // entries of lane sub of slot i: lcount[i * tpa + sub]
    int npad;
    int type_bits;
    int tpa;               // lanes per atom; entries grouped by tpa (see k_force)
    int built;             // 0: the list was never built -> no pairs
    // tiled layout (tiles.hpp): nbr holds 16-bit staged slots
    int tiled;
    int ntiles;            // tiles = workgroups of a tiled force launch
    int smax;              // largest staged stencil of any tile
    int wcap;              // 32-bit words per lane the list array holds (cap / tpa)
    const int4 *tiles;
    const int2 *truns;     // kTileRuns runs per tile (tiles.hpp)
    const uint32_t *keys;  // sorted cell keys of the last rebuild
    const int32_t *cstart;
    CellGrid grid;
};

struct PairView {
    int kind;
    int ntypes;
    int single;            // 1: one type pair (parameters uniform)
    int harmonic;          // Repulsion with n == 2 for every type pair
    const PairParam *tab;
    PairParam p0;          // tab[0] by value: a one-type kernel reads it from
                           // its arguments (scalar registers), not from memory
};

// Integrator stage fused into the force kernel.
enum Stage : int {
    kStageForceOnly = 0,   // Collection::set_forces(false): f only
    kStageForceA = 1,      // set_forces(true): f and a = f/m
    kStageVerletKick = 2,  // Verlet second half: a = f/m ; v += a dt/2
    kStageVerletKickDrift = 3,  // ... then the next step's drift + half kick
    kStageDampedKick = 4,  // Damped: a = f/m ; v += a dt c2
    kStageDampedKickDrift = 5,
    kStageSums = 6,        // energy / virial / stress only, nothing written
    kStageForceAdd = 7,    // NListed::set_forces: f += pair forces
};

struct StepCoef {
    // Verlet: dt, dt*dt/2, dt/2.  Damped: c1*dt, c2*dt*dt, dt*(c1-c2), c0, dt*c2
    double k0, k1, k2, k3, k4;
};

// Reduction scratch: one partial per workgroup, reduced by k_finalize.
struct ReduceBuf {
    double *part;                 // kRedWidth fields x pstride workgroups
    int pstride;                  // field stride (>= workgroups of any launch)
    double *out;                  // kRedWidth doubles (final sums)
    uint32_t *flag;               // flag word
    double skin;                  // skin test threshold
    uint32_t *hflag;              // device-visible pinned host copy of the flag word, or null
    hipEvent_t done;              // host side, or null: recorded on the launch stream right
                                  // after the force kernel, before its k_finalize (timing)
    uint32_t *cnt;                // or null: the skin-test reduction over several
    double *gpart;                // workgroups (counter, kFinBlocks x 3 results)
};
constexpr int kFinBlocks = 32;    // workgroups of the multi-workgroup skin reduction
constexpr int kRedWidth = 16;     // U, W, S[9], contacts, overlaps, top1, top2, nonfinite

hipError_t launch_force(int ndim, int precision, int stage, bool energy,
                        const AtomArrays &A, const ListView &Lv, const PairView &P,
                        const Box &box, const StepCoef &c, const ReduceBuf &R,
                        hipStream_t s);
// Dynamic LDS of the tiled force kernel for a stencil of smax staged slots
// (dimension, precision, pair kind, one pair type or several; no pair kind
// yet gets the largest staging).
size_t tiled_force_lds(int ndim, int prec, int kind, bool single, int smax);
// Tiled force launch (step_tiled.hip); same contract as launch_force.
hipError_t launch_force_tiled(int ndim, int precision, int stage, bool energy,
                              const AtomArrays &A, const ListView &Lv, const PairView &P,
                              const Box &box, const StepCoef &c, const ReduceBuf &R,
                              hipStream_t s);
// k_finalize over nparts workgroup partials.
hipError_t launch_finalize_rt(bool sums, bool skin, int nparts, const ReduceBuf &R,
                              hipStream_t s);
hipError_t launch_drift(int ndim, int stage, const AtomArrays &A, const StepCoef &c,
                        const ReduceBuf &R, hipStream_t s);
hipError_t launch_skin(int ndim, const AtomArrays &A, const ReduceBuf &R,
                       hipStream_t s);
// Atom observables: mode 0 -> out[0] = K(originv), out[1..D] = momentum,
// out[4] = mass ; mode 1 -> out[0] = sum |f|, out[13] = max |f|.
hipError_t launch_atom_sums(int ndim, int mode, const AtomArrays &A,
                            const double *originv, const ReduceBuf &R, hipStream_t s);
// Layout conversion between the caller's AoS order and the device slots.
// upload: dst component k of slot s = src[orig[s]*ndim + k]  (what: 0 pos,
// 1 v, 2 a, 3 f);  download: the inverse.
hipError_t launch_upload(int ndim, int what, const AtomArrays &A, const int32_t *orig,
                         const double *src, hipStream_t s);
hipError_t launch_download(int ndim, int what, const AtomArrays &A,
                           const int32_t *orig, double *dst, hipStream_t s);
int reduce_blocks(int n);
int force_blocks(int n, int tpa);  // workgroups of a force launch (also partials)

// ---------------------------------------------------------------- batch.hip
// Geometry of one system of a batch.
struct BatchGeom {
    Box box;
    CellGrid grid;
    int off;       // first atom of the system in the concatenated arrays
    int n;         // atoms
    int cell_off;  // first cell in the concatenated cell-start array
    int pad;
};

struct BatchArrays {
    int ntot;              // atoms over all systems
    const void *pos_in;    // PosStore<D>::T*
    void *pos_out;
    double *v[3], *a[3], *f[3];
    const double *m;
    const int32_t *type;   // pair-table type
    const int32_t *sys;    // system of each atom
    const double *diam;    // list diameter of each atom
    const void *xref;      // positions at each system's last rebuild
    void *xref_w;          // the same buffer, written by the rebuild
    double *disp;          // |x - xref| per atom (skin test input)
    const BatchGeom *geom; // per system
};

hipError_t launch_batch_force(int ndim, int precision, int stage, bool energy,
                              const BatchArrays &A, const ListView &Lv, const PairView &P,
                              const StepCoef &c, double *eatom, hipStream_t s);
hipError_t launch_batch_drift(int ndim, bool damped, const BatchArrays &A,
                              const StepCoef &c, hipStream_t s);
hipError_t launch_batch_disp(int ndim, const BatchArrays &A, hipStream_t s);
hipError_t launch_batch_skin(int nsys, const BatchArrays &A, double skin, uint32_t *flags,
                             uint32_t *gflag, hipStream_t s);
// Rebuilds the systems whose skin flag is set (all with force): worklist,
// sort, then the list build spread over slices of the listed systems.
// work: nsys + 1 ints (count, then system ids); wslice: nsys + 1 ints (each
// listed system's first build slice, then the total).  nent: per system the
// list entries of its last build (twice its pairs).
hipError_t launch_batch_rebuild(int ndim, int nsys, const BatchArrays &A, const void *pos,
                                const void *alt, int force, int count_rebuild, double skin,
                                int type_bits, int tpa, int cap, int npad, uint32_t *nbr,
                                int32_t *ncount, int32_t *keys, int32_t *cellidx,
                                int32_t *cstart, void *xs, double *dsort, uint32_t *flags,
                                uint32_t *which, unsigned long long *nent, uint32_t *gflag,
                                int32_t *work, int32_t *wslice, int sort_grid, int build_grid,
                                hipStream_t s);
hipError_t launch_batch_sys_sums(int nsys, const BatchArrays &A, const double *eatom,
                                 double *out, hipStream_t s);
constexpr int kBatchMaxCellsPerSystem = 4096;

// Sets the calling thread's parm_last_error() message; returns code.
int set_error(int code, const std::string &msg);

}  // namespace parm
