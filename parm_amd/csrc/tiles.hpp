// tiles.hpp -- row tiles and their LDS-staged neighbourhoods, shared by the
// list build (nlist.hip) and the force kernels (step_tiled.hip).
//
// After a rebuild the atoms are sorted by cell with x fastest, so the atoms
// of one row of cells (fixed y, z) are contiguous.  A tile is a run of at
// most kTileAtoms consecutive atoms of one row.  Every neighbour candidate of
// a tile's atoms lies in the tile's stencil: for each of the NR = 9 (2D: 3)
// rows at y-1..y+1, z-1..z+1, the x cells xa-1 .. xb+1 around the tile's cells
// [xa, xb].  Those cells are staged in LDS in (row, x, slot) order: the atoms
// of cell c of the stencil occupy staged slots off[c] .. off[c+1].  An atom in
// cell x sees exactly the staged range off[k*W + u-1] .. off[k*W + u+2] of
// stencil row k (u = x - xa + 1), one contiguous range per row.
//
// The list of a tiled layout stores the staged slot (16 bits) instead of the
// atom index: half the bytes, and the force kernel reads neighbour
// positions from LDS instead of gathering them from memory.
//
// When the stencil would wrap onto itself along x (xb - xa + 3 > nc0: the
// tile spans most of its row), the whole row is staged once in cell order
// instead ("full row": xa = 1, W = nc0, no periodic shift along x); an atom
// in the first or last cell then sees two ranges per stencil row, and the
// list build takes the minimum image along x explicitly.
//
// Atoms that are not list members (cell key = ncell, sorted last) get tail
// tiles with no stencil.
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace parm {

constexpr int kTileAtoms = 64;     // atoms per tile
constexpr int kTileWMax = 48;      // stencil cells along x per row, at most
constexpr int kTileMaxCells = 9 * kTileWMax;
constexpr int kTileMaxSlots = 65535;  // 16-bit list entries

// Runs: the stencil of a tile as at most kTileRuns contiguous atom ranges
// (per stencil row, the cells before and after the periodic wrap along x),
// in staged-slot order.  Run r of tile t is truns[t * kTileRuns + r] =
// (first atom, first staged slot); its length is the next run's first slot
// (the last run ends at the tile's S).
constexpr int kTileRuns = 18;

// Tile record: a0 first atom, cnt atoms, row (cy + nc1 * cz; -1: tail tile of
// non-members), S staged slots (-1: stencil does not fit; the system then
// uses the untiled layout).
using TileRec = int4;

// The stencil of one tile, built cooperatively in LDS.
struct TileLayout {
    int a0, cnt, row, S;
    int W, NR, xa, cy, cz, xfull;
    int off[kTileMaxCells + 1];  // first staged slot of stencil cell c
    int g0[kTileMaxCells];       // first atom of stencil cell c
    int cc[kTileMaxCells];       // atoms in stencil cell c
};

__device__ __forceinline__ int wrap_cell(int c, int n) {
    return c < 0 ? c + n : (c >= n ? c - n : c);
}

// Stencil rows along y and z: 3 when the grid has >= 3 cells there, else 1
// (a collapsed dimension holds every image in its one cell).
template <int D>
__device__ __forceinline__ int stencil_ny(const CellGrid &g) {
    return (D >= 2 && g.nc[1] >= 3) ? 3 : 1;
}
template <int D>
__device__ __forceinline__ int stencil_nz(const CellGrid &g) {
    return (D == 3 && g.nc[2] >= 3) ? 3 : 1;
}

// Cell index and periodic shift (in box lengths, -1/0/+1 per dimension) of
// stencil cell (k, u) of a tile at row (cy, cz) with first x cell xa.
template <int D>
__device__ __forceinline__ int stencil_cell(const CellGrid &g, int cy, int cz, int xa,
                                            int k, int u, int (&img)[3]) {
    const int ny = stencil_ny<D>(g);
    const int oy = ny == 3 ? (k % 3) - 1 : 0;
    const int oz = stencil_nz<D>(g) == 3 ? (k / ny) - 1 : 0;
    int x = 0;
    img[0] = img[1] = img[2] = 0;
    if (g.nc[0] >= 3) {
        x = xa - 1 + u;
        img[0] = x < 0 ? -1 : (x >= g.nc[0] ? 1 : 0);
        x = wrap_cell(x, g.nc[0]);
    }
    int y = 0, z = 0;
    if (D >= 2) {
        y = cy + oy;
        img[1] = y < 0 ? -1 : (y >= g.nc[1] ? 1 : 0);
        y = wrap_cell(y, g.nc[1]);
    }
    if (D == 3) {
        z = cz + oz;
        img[2] = z < 0 ? -1 : (z >= g.nc[2] ? 1 : 0);
        z = wrap_cell(z, g.nc[2]);
    }
    return (D == 3 ? (z * g.nc[1] + y) * g.nc[0] : y * g.nc[0]) + x;
}

// Stencil along x of a tile whose atoms span cells [xa, xb]: width W and
// first cell xa (rewritten to 1 for a full row); returns 1 for a full row.
__device__ __forceinline__ int stencil_x(const CellGrid &g, int &xa, int xb, int &W) {
    if (g.nc[0] < 3) {
        W = 1;
        return 0;
    }
    W = xb - xa + 3;
    if (W <= g.nc[0]) return 0;
    xa = 1;
    W = g.nc[0];
    return 1;
}

// First stencil column (u) past the periodic wrap along x, for a tile with
// first x cell xa and width W (W: no wrap).  Within each stencil row the
// columns before it and from it on are the row's two runs; each run is a
// contiguous range of cells with one periodic image.
__device__ __forceinline__ int stencil_ustar(const CellGrid &g, int xa, int W, int xfull) {
    if (g.nc[0] < 3 || xfull) return W;
    if (xa - 1 < 0) return 1;
    if (xa - 1 + W > g.nc[0]) return g.nc[0] - (xa - 1);
    return W;
}

// Periodic image (box lengths per dimension) of run r of a tile: stencil row
// r / 2, columns before (r even) or from (r odd) the wrap.  Zero for an
// empty run.
template <int D>
__device__ __forceinline__ void run_image(const TileLayout &L, const CellGrid &g, int r,
                                          int (&img)[3]) {
    img[0] = img[1] = img[2] = 0;
    const int ustar = stencil_ustar(g, L.xa, L.W, L.xfull);
    const int u0 = (r & 1) ? ustar : 0;
    if (r / 2 < L.NR && u0 < L.W) (void)stencil_cell<D>(g, L.cy, L.cz, L.xa, r / 2, u0, img);
}

// Staged ranges of stencil row k seen by an atom in x cell cx: [lo0, hi0)
// and [lo1, hi1) (the second one empty unless a full row wraps).
__device__ __forceinline__ void stencil_ranges(const TileLayout &L, const CellGrid &g,
                                               int k, int cx, int &lo0, int &hi0, int &lo1,
                                               int &hi1) {
    const int cb = k * L.W;
    lo1 = hi1 = 0;
    if (g.nc[0] < 3) {
        lo0 = L.off[cb];
        hi0 = L.off[cb + 1];
    } else if (!L.xfull) {
        const int u = cx - L.xa + 1;
        lo0 = L.off[cb + u - 1];
        hi0 = L.off[cb + u + 2];
    } else {
        const int n0 = g.nc[0];
        lo0 = L.off[cb + (cx > 0 ? cx - 1 : 0)];
        hi0 = L.off[cb + (cx < n0 - 1 ? cx + 2 : n0)];
        if (cx == 0) {
            lo1 = L.off[cb + n0 - 1];
            hi1 = L.off[cb + n0];
        } else if (cx == n0 - 1) {
            lo1 = L.off[cb];
            hi1 = L.off[cb + 1];
        }
    }
}

// Fills L for tile t (all threads of the block call it; ends with a barrier).
// Requires blockDim.x >= kWave.
template <int D>
__device__ void tile_layout(const TileRec *__restrict__ tiles, int t,
                            const uint32_t *__restrict__ keys,
                            const int32_t *__restrict__ cstart, const CellGrid &g,
                            TileLayout &L) {
    if (threadIdx.x == 0) {
        const TileRec T = tiles[t];
        L.a0 = T.x;
        L.cnt = T.y;
        L.row = T.z;
        L.S = T.w;
        if (T.z < 0 || T.y <= 0) {
            L.W = 0;
            L.NR = 0;
            L.xa = 0;
            L.xfull = 0;
            L.cy = L.cz = 0;
        } else {
            int xa = (int)(keys[T.x] % (uint32_t)g.nc[0]);
            const int xb = (int)(keys[T.x + T.y - 1] % (uint32_t)g.nc[0]);
            int W = 0;
            L.xfull = stencil_x(g, xa, xb, W);
            L.xa = xa;
            L.W = W;
            L.NR = stencil_ny<D>(g) * stencil_nz<D>(g);
            L.cy = D >= 2 ? T.z % g.nc[1] : 0;
            L.cz = D == 3 ? T.z / g.nc[1] : 0;
        }
    }
    __syncthreads();
    const int nc = L.NR * L.W;
    // per-cell counts, then an exclusive scan by the first wavefront
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        int img[3];
        const int cell = stencil_cell<D>(g, L.cy, L.cz, L.xa, c / L.W, c % L.W, img);
        const int s0 = cstart[cell];
        L.g0[c] = s0;
        L.cc[c] = cstart[cell + 1] - s0;
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        constexpr int kPer = (kTileMaxCells + kWave - 1) / kWave;
        int loc[kPer];
        int sum = 0;
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int c = lane * kPer + q;
            loc[q] = c < nc ? L.cc[c] : 0;
            sum += loc[q];
        }
        int incl = sum;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        int run = incl - sum;
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int c = lane * kPer + q;
            if (c < nc) L.off[c] = run;
            run += loc[q];
        }
        if (lane == kWave - 1) L.off[nc] = incl;
    }
    __syncthreads();
}

// Stencil cell holding staged slot s (0 <= s < off[nc]): the last c with
// off[c] <= s.
__device__ __forceinline__ int slot_cell(const TileLayout &L, int s) {
    int lo = 0, hi = L.NR * L.W;  // off[lo] <= s < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (L.off[mid] <= s)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// A tile's runs as the force kernels hold them in LDS: run r as (first atom -
// first slot, first slot), padded to kRunSlots entries with (0, INT_MAX), so
// that the atom of a staged slot is found by binary lifting with no bounds
// checks: five independent-of-the-data steps, each one LDS read.
constexpr int kRunSlots = 32;
static_assert(kRunSlots >= kTileRuns && (kRunSlots & (kRunSlots - 1)) == 0,
              "runs padded to a power of two");

__device__ __forceinline__ int2 run_entry(const int2 *truns, int t, int r) {
    if (r >= kTileRuns) return int2{0, 0x7fffffff};
    const int2 u = truns[(size_t)t * kTileRuns + r];
    return int2{u.x - u.y, u.y};
}

// Atoms of K staged slots s[q] (0 <= s[q] < S): s plus the offset of the last
// run whose first slot is <= s (empty runs share their first slot with the
// next run, which the lifting then takes), and that run's index.  The K
// liftings advance together, one LDS read each per step, so their reads are
// in flight at once.
template <int K>
__device__ __forceinline__ void slot_runs(const int2 *rs, const int (&s)[K], int (&r)[K],
                                          int (&j)[K]) {
    int lo[K], d[K];
    const int d0 = rs[0].x;
#pragma unroll
    for (int q = 0; q < K; q++) {
        lo[q] = 0;
        d[q] = d0;
    }
#pragma unroll
    for (int step = kRunSlots / 2; step > 0; step >>= 1) {
        int2 e[K];
#pragma unroll
        for (int q = 0; q < K; q++) e[q] = rs[lo[q] + step];
#pragma unroll
        for (int q = 0; q < K; q++) {
            const bool take = e[q].y <= s[q];
            lo[q] = take ? lo[q] + step : lo[q];
            d[q] = take ? e[q].x : d[q];
        }
    }
#pragma unroll
    for (int q = 0; q < K; q++) {
        j[q] = s[q] + d[q];
        r[q] = lo[q];
    }
}

template <int K>
__device__ __forceinline__ void slot_atoms(const int2 *rs, const int (&s)[K], int (&j)[K]) {
    int r[K];
    slot_runs<K>(rs, s, r, j);
}

__device__ __forceinline__ int slot_atom(const int2 *rs, int s) {
    const int sv[1] = {s};
    int j[1];
    slot_atoms<1>(rs, sv, j);
    return j[0];
}

// Position of list entry q (lane q % tpa, lane-local index m = q / tpa) in the
// 16-bit tiled list: word ((m / 2) * npad + i) * tpa + q % tpa holds lane
// entries m = 2w (low half) and 2w + 1 (high half).
__device__ __forceinline__ size_t tiled_entry(int q, int i, int tpa, int npad) {
    const int ts = tpa_shift(tpa);
    const int m = q >> ts;
    return (((((size_t)(m >> 1) * npad + i) << ts) + (q & (tpa - 1))) << 1) + (m & 1);
}

}  // namespace parm
