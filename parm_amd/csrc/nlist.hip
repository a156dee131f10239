// nlist.hip -- device Verlet-list rebuild for NeighborList::update_list
// (reference: src/trackers.cpp:19-85).
//
// The reference rebuild is an O(N^2) double loop over members testing
//   |diff(x_i, x_j)| < (d_i + d_j)/2 + skin.
// Here the same predicate, evaluated bit-identically, runs over a periodic
// cell grid (cell edge >= the largest list radius) after a stable LSD radix
// sort of the atoms by cell.  The sort also renumbers the device slots, so
// atoms that are close in space are close in memory for the force gathers.
//
// Pipeline (one rebuild):
//   k_cell_keys      key = cell of each member, ncell for non-members
//   radix passes     k_radix_hist -> k_radix_scan -> k_radix_scatter (8 bits)
//   k_permute        every per-slot array permuted into the new order
//   k_cell_bounds    first slot of each cell
//   k_build_list<D>  full (both directions) list, ELL layout nbr[k*npad + i]
//
// Tiled layout (tiles.hpp), used when every tile's stencil fits in LDS:
//   k_tile_rows      rows of cells -> tiles of <= kTileAtoms atoms
//   k_tile_sizes<D>  staged slots of each tile, their maximum
//   k_build_tiled<D> the tile's stencil staged in LDS, the same predicate
//                    evaluated from LDS, 16-bit staged-slot entries
#include <algorithm>
#include <type_traits>
#include <utility>

#include "common.hpp"
#include "kernels.hpp"
#include "tiles.hpp"

namespace parm {

// ------------------------------------------------------------ cell keys
template <int D>
__global__ __launch_bounds__(kBlock) void k_cell_keys(
    int n, const typename PosStore<D>::T *__restrict__ pos,
    const int32_t *__restrict__ rank, CellGrid g, uint32_t *__restrict__ keys,
    int32_t *__restrict__ vals) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    vals[i] = i;
    if (rank[i] < 0) {
        keys[i] = (uint32_t)g.ncell;
        return;
    }
    Vec<D> x = PosStore<D>::load(pos[i]);
    uint32_t key = 0;
#pragma unroll
    for (int k = D - 1; k >= 0; k--) {
        double s = x.c[k] * g.invL[k];
        s -= floor(s);
        int c = (int)(s * (double)g.nc[k]);
        c = c < 0 ? 0 : (c >= g.nc[k] ? g.nc[k] - 1 : c);
        key = key * (uint32_t)g.nc[k] + (uint32_t)c;
    }
    keys[i] = key;
}

// ------------------------------------------------------------ radix sort
// Stable LSD radix sort of (key, value) pairs, 8 bits per pass.  Each block
// owns a tile of kRadixTile consecutive items.  hist[d * nblocks + b] counts
// digit d in tile b; its exclusive scan is the output offset of (d, b), which
// keeps equal digits in input order across tiles.  Within a tile the scatter
// ranks equal digits in input order with 64-lane ballots.
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kBlock * kRadixItems;

__global__ __launch_bounds__(kBlock) void k_radix_hist(
    int n, int shift, const uint32_t *__restrict__ keys,
    uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kRadixTile;
    for (int t = 0; t < kRadixItems; t++) {
        int i = base + t * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of one value per thread over a workgroup of NT threads;
// total = the sum over the workgroup.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t &total) {
    __shared__ uint32_t wsum[NT / kWave];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int q = 0; q < NT / kWave; q++) {
        const uint32_t ws = wsum[q];
        if (q < w) before += ws;
        total += ws;
    }
    __syncthreads();
    return before + incl - v;
}

// One workgroup per digit d: exclusive scan of the digit's per-tile counts
// hist[d * nb + b] in place (each thread a contiguous chunk), and
// dtot[d] = the digit's total.  The scatter adds the digits' exclusive scan.
__global__ __launch_bounds__(kBlock) void k_radix_scan(int nb, uint32_t *__restrict__ hist,
                                                     uint32_t *__restrict__ dtot) {
    uint32_t *h = hist + (size_t)blockIdx.x * nb;
    const int per = (nb + kBlock - 1) / kBlock;
    const int lo = min(nb, (int)threadIdx.x * per), hi = min(nb, lo + per);
    uint32_t s = 0;
    for (int i = lo; i < hi; i++) s += h[i];
    uint32_t total = 0;
    uint32_t run = block_excl_scan_u32<kBlock>(s, total);
    for (int i = lo; i < hi; i++) {
        const uint32_t v = h[i];
        h[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    int n, int shift, const uint32_t *__restrict__ keys_in,
    const int32_t *__restrict__ vals_in, const uint32_t *__restrict__ offs,
    const uint32_t *__restrict__ dtot, uint32_t *__restrict__ keys_out,
    int32_t *__restrict__ vals_out) {
    static_assert(kBlock == 256, "one thread per digit");
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[kBlock / kWave][256];
    const int t = threadIdx.x;
    const int w = t / kWave;
    const int lane = t % kWave;
    // first output slot of digit t in this tile: the totals of the smaller
    // digits plus the counts of digit t in the earlier tiles
    uint32_t all = 0;
    base[t] = block_excl_scan_u32<kBlock>(dtot[t], all) + offs[t * gridDim.x + blockIdx.x];
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int tile = blockIdx.x * kRadixTile;
    for (int c = 0; c < kRadixItems; c++) {
#pragma unroll
        for (int q = 0; q < kBlock / kWave; q++) wcnt[q][t] = 0;
        __syncthreads();
        int i = tile + c * kBlock + t;
        bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t d = (key >> shift) & 255u;
        unsigned long long eq = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            unsigned long long bb = __ballot(bit);
            eq &= bit ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(eq & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(eq);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d];
            for (int q = 0; q < w; q++) off += wcnt[q][d];
            off += rank;
            keys_out[off] = key;
            vals_out[off] = vals_in[i];
        }
        __syncthreads();
        // advance the per-digit base past this chunk (thread t owns digit t)
        uint32_t tot = 0;
#pragma unroll
        for (int q = 0; q < kBlock / kWave; q++) tot += wcnt[q][t];
        base[t] += tot;
        __syncthreads();
    }
}

// --------------------------------------------------------------- gathers
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(int n, const int32_t *__restrict__ perm,
                                                   const T *__restrict__ src,
                                                   T *__restrict__ dst) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_permute(int n, const int32_t *__restrict__ perm,
                                                    PermuteSet P) {
    using PT = typename PosStore<D>::T;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    // every load first, then every store: the gathers are in flight together
    const PT x0 = static_cast<const PT *>(P.pos_src[0])[p];
    const PT x1 = static_cast<const PT *>(P.pos_src[1])[p];
    double d[kPermD];
    int32_t q[kPermI];
#pragma unroll
    for (int c = 0; c < kPermD; c++)
        if (c < P.nd) d[c] = P.d_src[c][p];
#pragma unroll
    for (int c = 0; c < kPermI; c++)
        if (c < P.ni) q[c] = P.i_src[c][p];
    static_cast<PT *>(P.pos_dst[0])[i] = x0;
    static_cast<PT *>(P.pos_dst[1])[i] = x1;
#pragma unroll
    for (int c = 0; c < kPermD; c++)
        if (c < P.nd) P.d_dst[c][i] = d[c];
#pragma unroll
    for (int c = 0; c < kPermI; c++)
        if (c < P.ni) P.i_dst[c][i] = q[c];
}

// ---------------------------------------------------------- cell bounds
__global__ __launch_bounds__(kBlock) void k_cell_bounds(
    int n, int ncell, const uint32_t *__restrict__ keys,
    int32_t *__restrict__ start) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i > n) return;
    long prev = (i == 0) ? -1 : (long)keys[i - 1];
    long cur = (i == n) ? (long)ncell : (long)keys[i];
    if (i == n) {
        for (long c = prev + 1; c <= ncell; c++) start[c] = n;
        return;
    }
    for (long c = prev + 1; c <= cur && c <= ncell; c++) start[c] = i;
}

// ----------------------------------------------------------- list build
// G lanes per slot.  The candidates of slot i are the slots of its stencil,
// taken as rows of three x-adjacent cells, which are contiguous slot ranges
// (two ranges where the row wraps around the box).  In each round the G lanes
// test G consecutive candidates and the passing ones are appended in
// candidate order (ballot + prefix popcount), so the list order is
// deterministic.  Entry q of slot i goes to ((q / tpa) * npad + i) * tpa +
// q % tpa: groups of tpa consecutive entries, the layout k_force reads.
//
// The predicate is the reference's, bit for bit:
//   r = diff(x_i, x_j) (remainder per component), d2 = (x^2 + y^2) + z^2,
//   in list  <=>  sqrt(d2) < (d_i + d_j)/2 + skin   (correctly rounded sqrt).
// sqrt is only evaluated when d2 is within 1e-14 (relative) of c^2; outside
// that band the comparison with c^2 decides it with margin to spare.
// Is the pair (original atoms oi, oj) excluded (NeighborList::ignore)?
// Binary search of oi's sorted partner list.
__device__ __forceinline__ bool excluded(const ExclView &ex, int oi, int oj) {
    int lo = ex.off[oi], hi = ex.off[oi + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = ex.part[mid];
        if (v == oj) return true;
        if (v < oj)
            lo = mid + 1;
        else
            hi = mid;
    }
    return false;
}

template <int D, int G, int U>
__global__ __launch_bounds__(kBlock) void k_build_list(
    int n, const typename PosStore<D>::T *__restrict__ pos,
    const typename PosStore<D>::T *__restrict__ xw,
    const int32_t *__restrict__ rank, const double *__restrict__ diam,
    const int32_t *__restrict__ type, const uint32_t *__restrict__ keys,
    const int32_t *__restrict__ cstart, CellGrid g, Box box, double skin, double dmax,
    int type_bits, int tpa, int cap, int npad, uint32_t *__restrict__ nbr,
    int32_t *__restrict__ ncount, uint16_t *__restrict__ lcount,
    typename PosStore<D>::T *__restrict__ xref,
    unsigned long long *__restrict__ part /* per workgroup: sum, max */,
    uint32_t *__restrict__ flag, ExclView ex) {
    __shared__ unsigned long long red[kBlock / kWave][2];
    const int gt = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int i = gt / G;
    const int sub = gt % G;
    const int lane = threadIdx.x % kWave;
    const int gbase = lane - sub;  // first lane of this slot's group
    const unsigned long long below = (sub == 0) ? 0ull : ((1ull << sub) - 1ull);
    unsigned long long mycnt = 0;
    if (i < n) {
        const typename PosStore<D>::T pi = pos[i];
        if (sub == 0) xref[i] = pi;
        if (rank[i] < 0) {
            if (sub == 0) {
                ncount[i] = 0;
                for (int l = 0; l < tpa; l++) lcount[(size_t)i * tpa + l] = 0;
            }
        } else {
            const Vec<D> xi = PosStore<D>::load(pi);
            const Vec<D> wi = PosStore<D>::load(xw[i]);
            const double di = diam[i];
            const int oi = ex.any ? ex.orig[i] : 0;
            uint32_t key = keys[i];
            int ci[3] = {0, 0, 0};
#pragma unroll
            for (int k = 0; k < D; k++) {
                ci[k] = (int)(key % (uint32_t)g.nc[k]);
                key /= (uint32_t)g.nc[k];
            }
            const bool wide0 = g.nc[0] >= 3;
            const int ny = (D >= 2 && g.nc[1] >= 3) ? 3 : 1;
            const int nz = (D == 3 && g.nc[2] >= 3) ? 3 : 1;
            const int nrows = ny * nz;
            // The x cells of each row as at most two contiguous cell ranges
            // (the second one across the periodic boundary, shifted by -L or
            // +L).  Lane `sub` of the group loads the slot bounds of rows
            // sub and sub + G, so the group's loads are in flight together;
            // the loop below reads them back with shuffles.
            int xa0 = 0, xb0 = 0, xa1 = 0, xb1 = -1;
            double shx1 = 0.0;
            if (wide0) {
                if (ci[0] == 0) {
                    xa0 = 0; xb0 = 1; xa1 = g.nc[0] - 1; xb1 = g.nc[0] - 1;
                    shx1 = -box.L[0];
                } else if (ci[0] == g.nc[0] - 1) {
                    xa0 = ci[0] - 1; xb0 = ci[0]; xa1 = 0; xb1 = 0;
                    shx1 = box.L[0];
                } else {
                    xa0 = ci[0] - 1; xb0 = ci[0] + 1;
                }
            }
            auto row_base = [&](int r, double *sh) {
                const int oy = (ny == 3) ? (r % 3) - 1 : 0;
                const int oz = (nz == 3) ? (r / ny) - 1 : 0;
                int q1 = ci[1] + oy, q2 = ci[2] + oz;
                sh[1] = 0.0;
                sh[2] = 0.0;
                if (D >= 2) {
                    if (q1 < 0) { q1 += g.nc[1]; sh[1] = -box.L[1]; }
                    else if (q1 >= g.nc[1]) { q1 -= g.nc[1]; sh[1] = box.L[1]; }
                }
                if (D == 3) {
                    if (q2 < 0) { q2 += g.nc[2]; sh[2] = -box.L[2]; }
                    else if (q2 >= g.nc[2]) { q2 -= g.nc[2]; sh[2] = box.L[2]; }
                }
                return (D == 3 ? q2 * g.nc[1] + q1 : q1) * g.nc[0];
            };
            int bnd[2][4];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int r = sub + t * G;
                bnd[t][0] = bnd[t][1] = bnd[t][2] = bnd[t][3] = 0;
                if (r < nrows) {
                    double shd[3];
                    const int rowc = row_base(r, shd);
                    bnd[t][0] = cstart[rowc + xa0];
                    bnd[t][1] = cstart[rowc + xb0 + 1];
                    if (xb1 >= 0) {
                        bnd[t][2] = cstart[rowc + xa1];
                        bnd[t][3] = cstart[rowc + xb1 + 1];
                    }
                }
            }
            int cnt = 0;
            for (int r = 0; r < nrows; r++) {
                double sh[3];
                (void)row_base(r, sh);
                const int t = r / G, owner = gbase + (r % G);
                int rb[4];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    rb[k] = __shfl(t == 0 ? bnd[0][k] : bnd[1][k], owner);
                for (int rr = 0; rr < 2; rr++) {
                    const int a = rb[2 * rr], b = rb[2 * rr + 1];
                    sh[0] = rr == 0 ? 0.0 : shx1;
                    // U candidates per lane per round, their loads issued
                    // together; appended in candidate order
                    for (int base = a; base < b; base += G * U) {
                        typename PosStore<D>::T pw[U];
                        double dj[U];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const int j = base + u * G + sub;
                            const int jj = j < b ? j : i;
                            pw[u] = xw[jj];
                            dj[u] = diam[jj];
                        }
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const int j = base + u * G + sub;
                            bool pass = false;
                            if (j < b && j != i) {
                                // |r|^2 from the wrapped copies and the row's
                                // periodic shift: equal to the exact d2 up
                                // to the rounding of the copies, far below
                                // the 1e-6 band; inside the band the
                                // reference's predicate decides on the
                                // unwrapped positions.
                                const Vec<D> wj = PosStore<D>::load(pw[u]);
                                double d2w = 0.0;
#pragma unroll
                                for (int k = 0; k < D; k++) {
                                    double dk = wi.c[k] - (wj.c[k] + sh[k]);
                                    // a collapsed dimension (one cell) keeps
                                    // every image: take the nearest one
                                    if (g.nc[k] < 3) dk -= box.L[k] * rint(dk * box.invL[k]);
                                    d2w = fma(dk, dk, d2w);
                                }
                                const double c = (di + dj[u]) / 2 + skin;
                                const double c2 = c * c;
                                if (d2w < c2 * (1.0 - 1e-6))
                                    pass = true;
                                else if (d2w <= c2 * (1.0 + 1e-6))
                                    pass = in_list<D>(xi, PosStore<D>::load(pos[j]), di,
                                                      dj[u], box, skin);
                                // ignored pairs never enter the list
                                // (src/trackers.cpp:64)
                                if (ex.any && pass) pass = !excluded(ex, oi, ex.orig[j]);
                            }
                            const unsigned long long m =
                                (__ballot(pass) >> gbase) &
                                ((G == 64) ? ~0ull : ((1ull << G) - 1ull));
                            if (pass) {
                                const int q = cnt + __popcll(m & below);
                                if (q < cap)
                                    nbr[untiled_entry(q, i, tpa, npad)] =
                                        ((uint32_t)j << type_bits) | (uint32_t)type[j];
                            }
                            cnt += __popcll(m);
                        }
                    }
                }
            }
            if (sub == 0) {
                const int cc = cnt < cap ? cnt : cap;
                ncount[i] = cc;
                // entry q is lane q % tpa's entry q / tpa
                for (int l = 0; l < tpa; l++)
                    lcount[(size_t)i * tpa + l] = (uint16_t)(cc > l ? (cc - l + tpa - 1) / tpa : 0);
                if (cnt > cap) atomicOr(flag, kFlagOverflow);
                mycnt = (unsigned long long)cnt;
            }
        }
    }
    // per-workgroup sum and max of the counts (no same-address atomics)
    unsigned long long sum = mycnt, mx = mycnt;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off);
        const unsigned long long o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
    }
    if (lane == 0) {
        red[threadIdx.x / kWave][0] = sum;
        red[threadIdx.x / kWave][1] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s2 = 0, m2 = 0;
        for (int q = 0; q < kBlock / kWave; q++) {
            s2 += red[q][0];
            m2 = red[q][1] > m2 ? red[q][1] : m2;
        }
        part[2 * (size_t)blockIdx.x] = s2;
        part[2 * (size_t)blockIdx.x + 1] = m2;
    }
}

// ------------------------------------------------------------ row tiles
// One workgroup of 1024 threads.  Row r (cells r*nc0 .. r*nc0 + nc0 - 1)
// holds atoms cstart[r*nc0] .. cstart[(r+1)*nc0]; it is cut into
// ceil(n_r / kTileAtoms) tiles in order.  Non-members (slots from
// cstart[ncell] on) get tail tiles.  info[0] = number of tiles (0 and
// info[2] |= 1 if they exceed cap_tiles).
constexpr int kTileScan = 1024;

__device__ __forceinline__ int block_excl_scan(int v, int &total) {
    __shared__ int wsum[kTileScan / kWave];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    int incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int q = 0; q < kTileScan / kWave; q++) {
        const int ws = wsum[q];
        if (q < w) before += ws;
        total += ws;
    }
    __syncthreads();
    return before + incl - v;
}

__global__ __launch_bounds__(kTileScan) void k_tile_rows(
    int n, int nrows, int nc0, int ncell, const int32_t *__restrict__ cstart,
    int cap_tiles, TileRec *__restrict__ tiles, int *__restrict__ info) {
    int base = 0;
    for (int r0 = 0; r0 < nrows; r0 += kTileScan) {
        const int r = r0 + (int)threadIdx.x;
        int a = 0, cnt = 0, nt = 0;
        if (r < nrows) {
            a = cstart[r * nc0];
            cnt = cstart[(r + 1) * nc0] - a;
            nt = (cnt + kTileAtoms - 1) / kTileAtoms;
        }
        int total = 0;
        const int off = base + block_excl_scan(nt, total);
        for (int q = 0; q < nt; q++) {
            const int idx = off + q;
            if (idx < cap_tiles)
                tiles[idx] = TileRec{a + q * kTileAtoms,
                                     min(kTileAtoms, cnt - q * kTileAtoms), r, 0};
        }
        base += total;
    }
    const int m0 = cstart[ncell];
    const int ntail = (n - m0 + kTileAtoms - 1) / kTileAtoms;
    for (int q = (int)threadIdx.x; q < ntail; q += kTileScan) {
        const int idx = base + q;
        if (idx < cap_tiles)
            tiles[idx] = TileRec{m0 + q * kTileAtoms, min(kTileAtoms, n - m0 - q * kTileAtoms),
                                 -1, 0};
    }
    if (threadIdx.x == 0) {
        const int nt = base + ntail;
        info[0] = nt <= cap_tiles ? nt : 0;
        info[1] = 0;
        info[2] = nt <= cap_tiles ? 0 : 1;
    }
}

// One thread per tile: stencil width and staged slots.  A tile whose stencil
// is wider than kTileWMax cells or holds more than kTileMaxSlots atoms marks
// info[2]; the system then uses the untiled layout.
template <int D>
__global__ __launch_bounds__(kBlock) void k_tile_sizes(int cap_tiles, CellGrid g,
                                                       const uint32_t *__restrict__ keys,
                                                       const int32_t *__restrict__ cstart,
                                                       TileRec *__restrict__ tiles,
                                                       int2 *__restrict__ truns,
                                                       int *__restrict__ info) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int nt = info[0];
    int S = 0, bad = 0;
    if (t < nt && t < cap_tiles) {
        TileRec T = tiles[t];
        int2 *runs = truns + (size_t)t * kTileRuns;
        int nr = 0;
        if (T.z >= 0 && T.y > 0) {
            int xa = (int)(keys[T.x] % (uint32_t)g.nc[0]);
            const int xb = (int)(keys[T.x + T.y - 1] % (uint32_t)g.nc[0]);
            int W = 0;
            const int xfull = stencil_x(g, xa, xb, W);
            if (W > kTileWMax) {
                bad = 1;
            } else {
                const int cy = D >= 2 ? T.z % g.nc[1] : 0;
                const int cz = D == 3 ? T.z / g.nc[1] : 0;
                const int NR = stencil_ny<D>(g) * stencil_nz<D>(g);
                // per stencil row: the cells before / after the wrap along x
                const int ustar = stencil_ustar(g, xa, W, xfull);
                for (int k = 0; k < NR; k++) {
                    for (int h = 0; h < 2; h++) {
                        const int u0 = h == 0 ? 0 : ustar, u1 = h == 0 ? ustar : W;
                        int g0 = 0, len = 0;
                        if (u1 > u0) {
                            int img[3];
                            const int c0 = stencil_cell<D>(g, cy, cz, xa, k, u0, img);
                            const int c1 = stencil_cell<D>(g, cy, cz, xa, k, u1 - 1, img);
                            g0 = cstart[c0];
                            len = cstart[c1 + 1] - g0;
                        }
                        runs[nr++] = int2{g0, S};
                        S += len;
                    }
                }
                if (S > kTileMaxSlots) bad = 1;
            }
        }
        for (; nr < kTileRuns; nr++) runs[nr] = int2{0, S};
        T.w = bad ? -1 : S;
        tiles[t] = T;
    }
    // one atomic per wavefront
    int smax = bad ? 0 : S;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        smax = max(smax, __shfl_xor(smax, o));
        bad |= __shfl_xor(bad, o);
    }
    if (threadIdx.x % kWave == 0) {
        if (smax > 0) atomicMax(info + 1, smax);
        if (bad) atomicOr(info + 2, 1);
    }
}

// Inclusive prefix sum of v over the G consecutive lanes of a group (lane
// `sub` of the group; glast = the wave lane of the group's last lane) and
// the group total.  For G = 4 the groups are the DPP quads: quad_perm moves
// (VALU, no LDS crossbar) shift the values up by 1 and 2 lanes and
// broadcast lane 3.
template <int G>
__device__ __forceinline__ void group_scan(int v, int sub, int glast, int &incl, int &tot) {
    if constexpr (G == 4) {
        incl = v;
        const int y1 = __builtin_amdgcn_update_dpp(0, incl, 0x90, 0xF, 0xF, false);  // [0,0,1,2]
        if (sub >= 1) incl += y1;
        const int y2 = __builtin_amdgcn_update_dpp(0, incl, 0x44, 0xF, 0xF, false);  // [0,1,0,1]
        if (sub >= 2) incl += y2;
        tot = __builtin_amdgcn_update_dpp(0, incl, 0xFF, 0xF, 0xF, false);  // [3,3,3,3]
    } else {
        incl = v;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (sub >= o) incl += y;
        }
        tot = __shfl(incl, glast);
    }
}

// Tiled list build: one workgroup per tile, G lanes per atom.  The tile's
// stencil is staged in LDS as FP32 offsets from the tile's first atom (its
// wrapped position), taken from the wrapped positions shifted by the stencil
// cell's periodic image, with the list diameter in the spare lane.  Each atom
// then scans one contiguous staged range per stencil row (two for an end cell
// of a full-row stencil).
//
// The FP32 test is only a prefilter.  With M the largest offset magnitude in
// the tile (and L along folded dimensions), every FP32 displacement component
// is within e = 2^-23 (M + L + c) of the exact one, so |d2_32 - d2| <=
// 2 sqrt(D) c e + 4 ulp(c^2); the tolerance is four times that.  Candidates
// farther than the tolerance from the cutoff are decided by the FP32 test;
// the others by the reference's exact predicate (remainder minimum image,
// sqrt(d2) < (d_i + d_j)/2 + skin) on the unwrapped FP64 positions, so the
// pair set is the reference's exactly.  Entries are staged slots.
// G lanes per atom.  In each round a lane tests U candidates; one prefix
// scan over the atom's G lanes then places the round's passing candidates in
// candidate-lane order, so entry q of the atom goes to the force kernel's
// lane q % tpa (its entry q / tpa) and every lane holds ceil or floor of
// count / tpa entries.  lcount[i * tpa + l] is lane l's count.
// OUT32: write untiled entries ((j << type_bits) | type_j, k_build_list's
// layout) instead of staged slots, for systems whose force kernel runs on the
// untiled list.  UNI: every member has the same list diameter; the staged
// slot then holds the offsets only (one cutoff for every candidate).
constexpr int kBuildPad = 64;  // staged slots of padding after the stencil

template <int D, bool UNI>
struct BuildSlot {
    static constexpr int W = UNI ? D : D + 1;  // floats per slot
    float v[W];
};

template <int D, int G, int U, bool OUT32, bool UNI>
__global__ __launch_bounds__(kWave *G) void k_build_tiled(
    const TileRec *__restrict__ tiles, const int2 *__restrict__ truns, int smax, const typename PosStore<D>::T *__restrict__ pos,
    const typename PosStore<D>::T *__restrict__ xw, const int32_t *__restrict__ rank,
    const double *__restrict__ diam, const int32_t *__restrict__ type, int type_bits,
    const uint32_t *__restrict__ keys, const int32_t *__restrict__ cstart, CellGrid g, Box box,
    double skin, double dmax, int tpa, int cap, int npad, int off24, void *__restrict__ nbr,
    int32_t *__restrict__ ncount, uint16_t *__restrict__ lcount,
    typename PosStore<D>::T *__restrict__ xref, unsigned long long *__restrict__ part,
    uint32_t *__restrict__ flag, const int *__restrict__ spec, ExclView ex) {
    constexpr int NT = kWave * G;
    using Slot = BuildSlot<D, UNI>;
    extern __shared__ float st_raw[];
    // smax staged slots plus one round of padding (kBuildPad >= G * U): a
    // round reads its U slots per lane unclamped, the ones past the range's
    // end are masked out
    Slot *st_f = reinterpret_cast<Slot *>(st_raw);
    uint32_t *st_e = reinterpret_cast<uint32_t *>(st_f + smax + kBuildPad);  // OUT32 entries
    static_assert(G * U <= kBuildPad, "a round stays inside the padded stage");
    static_assert(U <= 32, "a round's candidates are the bits of one word");
    constexpr int kBatch = U < 4 ? U : 4;  // slots read and tested together
    static_assert(U % kBatch == 0, "a round is whole batches");
    __shared__ TileLayout L;
    __shared__ int2 runs[kRunSlots];  // run_entry form (tiles.hpp)
    __shared__ double rimg[kTileRuns][3];
    __shared__ unsigned long long red[G][2];
    __shared__ int mbits;
    // tiles spread over the XCDs by the plan's tile count (a speculative
    // grid is a bound: its workgroups past the count have no tile)
    const int nt = spec ? spec[0] : (int)gridDim.x;
    const int t = (int)blockIdx.x < nt ? xcd_block(blockIdx.x, nt) : nt;
    if (spec) {
        // speculative launch (grid and smax are bounds): the decision is the
        // same for every thread of the workgroup
        bool skip = t >= nt;
        if (!skip) {
            const int S0 = tiles[t].w;
            if (S0 < 0 || S0 > smax) {
                skip = true;
                if (threadIdx.x == 0) atomicOr(flag, kFlagReplan);
            }
        }
        if (skip) {
            if (threadIdx.x == 0) {
                part[2 * (size_t)blockIdx.x] = 0;
                part[2 * (size_t)blockIdx.x + 1] = 0;
            }
            return;
        }
    }
    if (threadIdx.x == 0) mbits = 0;
    tile_layout<D>(tiles, t, keys, cstart, g, L);
    const int S = L.off[L.NR * L.W];
    // the stencil's runs (contiguous atoms, one periodic image each)
    if (threadIdx.x < kRunSlots) runs[threadIdx.x] = run_entry(truns, t, threadIdx.x);
    if (threadIdx.x < kTileRuns) {
        const int r = threadIdx.x;
        int img[3];
        run_image<D>(L, g, r, img);
#pragma unroll
        for (int k = 0; k < 3; k++) rimg[r][k] = k < D ? img[k] * box.L[k] : 0.0;
    }
    Vec<D> wo;
#pragma unroll
    for (int k = 0; k < D; k++) wo.c[k] = 0.0;
    if (S > 0) wo = PosStore<D>::load(xw[L.a0]);
    __syncthreads();
    float mloc = 0.0f;
    // kStage slots per thread at a time: their atoms and runs found by binary
    // lifting (the kStage liftings advancing together), then all their loads
    // in flight together, then the stores
    constexpr int kStage = 4;
    for (int s0 = threadIdx.x; s0 < S; s0 += kStage * NT) {
        int sb[kStage], rb[kStage], jb[kStage];
#pragma unroll
        for (int q = 0; q < kStage; q++) sb[q] = s0 + q * NT;
        slot_runs<kStage>(runs, sb, rb, jb);
        typename PosStore<D>::T wb[kStage];
        float db[kStage];
        int tb[kStage];
#pragma unroll
        for (int q = 0; q < kStage; q++) {
            if (sb[q] < S) {
                wb[q] = xw[jb[q]];
                if constexpr (!UNI) db[q] = (float)diam[jb[q]];
                if (OUT32) tb[q] = type[jb[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < kStage; q++) {
            const int s = sb[q];
            if (s >= S) break;
            const Vec<D> w = PosStore<D>::load(wb[q]);
            Slot sl;
#pragma unroll
            for (int k = 0; k < D; k++) {
                sl.v[k] = (float)((w.c[k] + rimg[rb[q]][k]) - wo.c[k]);
                mloc = fmaxf(mloc, fabsf(sl.v[k]));
            }
            if constexpr (!UNI) sl.v[D] = db[q];
            st_f[s] = sl;
            if (OUT32) st_e[s] = ((uint32_t)jb[q] << type_bits) | (uint32_t)tb[q];
        }
    }
    if (mloc > 0.0f) atomicMax(&mbits, __float_as_int(mloc));  // >= 0: int order
    __syncthreads();
    const int li = threadIdx.x / G, sub = threadIdx.x % G;
    const int lane = threadIdx.x % kWave;
    const int glast = lane - sub + G - 1;  // last lane of this atom's group
    unsigned long long mysum = 0, myreq = 0;
    if (li < L.cnt) {
        const int i = L.a0 + li;
        const typename PosStore<D>::T pi = pos[i];
        if (sub == 0) xref[i] = pi;
        if (L.row < 0 || rank[i] < 0) {
            if (sub == 0) {
                ncount[i] = 0;
                for (int l = 0; l < tpa; l++) lcount[(size_t)i * tpa + l] = 0;
            }
        } else {
            const Vec<D> xi = PosStore<D>::load(pi);
            const Vec<D> wi = PosStore<D>::load(xw[i]);
            const double di = diam[i];
            const float dif = (float)di;
            const int oi = ex.any ? ex.orig[i] : 0;
            const int cx = (int)(keys[i] % (uint32_t)g.nc[0]);
            // along x the minimum image is explicit for a collapsed
            // dimension or a full-row stencil (no staged shift)
            const bool fold0 = g.nc[0] < 3 || L.xfull;
            // Lf: the box edge along the folded dimensions, 0 along the
            // others (their fold then subtracts nothing: branch-free)
            bool foldk[3];
            float Lf[3], iLf[3], ri[3];
            double Lfold = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) {
                foldk[k] = k == 0 ? fold0 : g.nc[k] < 3;
                Lf[k] = foldk[k] ? (float)box.L[k] : 0.0f;
                iLf[k] = (float)box.invL[k];
                ri[k] = (float)(wi.c[k] - wo.c[k]);
                if (foldk[k]) Lfold = fmax(Lfold, box.L[k]);
            }
            // prefilter tolerance on d2 (see above)
            const double cmax = dmax + skin;
            const double e = 0x1p-23 * ((double)__int_as_float(mbits) + Lfold + cmax);
            const float tol = (float)(4.0 * (2.0 * sqrt((double)D) * cmax * e +
                                             4.0 * 0x1p-24 * cmax * cmax));
            const float skf = (float)skin;
            // atom i's own staged slot (centre stencil row, its cell)
            int self = -1;
            {
                const int kc = (stencil_ny<D>(g) == 3 ? 1 : 0) +
                               (stencil_nz<D>(g) == 3 ? stencil_ny<D>(g) : 0);
                const int uc = g.nc[0] < 3 ? 0 : (L.xfull ? cx : cx - L.xa + 1);
                const int cc = kc * L.W + uc;
                self = L.off[cc] + (i - L.g0[cc]);
            }
            int cnt = 0;  // the atom's entries so far (same in all its lanes)
            // off24 (the list spans under 4 GiB and an entry row of npad
            // atoms under 2^24 entries): entry q's element index in 32-bit
            // arithmetic, lane row (q >> ts) or row pair (q >> (ts + 1)) times
            // the row's entries by a 24-bit multiply, plus this atom's and
            // the lane's place in it (the 2-byte layout packs two rows per
            // 32-bit word: tiles.hpp tiled_entry)
            const int ts = tpa_shift(tpa);
            const uint32_t rowe = OUT32 ? ((uint32_t)npad << ts) : ((uint32_t)npad << (ts + 1));
            const uint32_t ie = OUT32 ? ((uint32_t)i << ts) : ((uint32_t)i << (ts + 1));
            // The candidate scan, compiled with and without FOLD (a minimum
            // image taken explicitly along some dimension).  With UNI the
            // one cutoff is the same FP32 value the per-pair formula gives.
            const bool anyfold = foldk[0] || (D >= 2 && foldk[1]) || (D == 3 && foldk[2]);
            const float cu = fmaf(0.5f, dif + dif, skf);
            const float c2u = cu * cu;
            const float c2lou = c2u - tol, c2hiu = c2u + tol;
            auto scan = [&](auto FOLDC) {
                constexpr bool FOLD = decltype(FOLDC)::value;
                for (int k = 0; k < L.NR; k++) {
                    int lo0, hi0, lo1, hi1;
                    stencil_ranges(L, g, k, cx, lo0, hi0, lo1, hi1);
                    // the second range (an end cell of a full-row stencil)
                    // only when some lane of the wavefront has one
                    const int nh = __ballot(hi1 > lo1) != 0ull ? 2 : 1;
                    for (int h = 0; h < nh; h++) {
                        const int lo = h ? lo1 : lo0, hi = h ? hi1 : hi0;
                        for (int base = lo; base < hi; base += G * U) {
                            // candidate u of this lane is slot base + u G + sub;
                            // `valid`: the ones below hi that are not the atom
                            // itself (one mask per round, not a test per slot)
                            const int first = base + sub;
                            const int nv = min(U, max(0, (hi - first + G - 1) / G));
                            uint32_t valid = nv >= 32 ? ~0u : ((1u << nv) - 1u);
                            {
                                const int ds = self - first;
                                if (ds >= 0 && (ds % G) == 0 && ds / G < U)
                                    valid &= ~(1u << (ds / G));
                            }
                            // bit u: d2 below the cutoff by more than the
                            // tolerance (pbits), below the cutoff plus the
                            // tolerance (ubits); filled from the top so that bit
                            // u is candidate u.  Each bit is the sign of d2 minus
                            // the bound (exact in sign: the difference of two
                            // finite floats rounds to zero only when they are
                            // equal), shifted in by one funnel shift.  At
                            // d2 = c2 + tol the candidate counts as out: the
                            // tolerance is four times the FP32 error bound.
                            // The slots are read and tested kBatch at a time (the
                            // last batch first), the scheduler kept from hoisting
                            // the next batch's reads above this one's tests: the
                            // round's registers stay those of one batch.
                            uint32_t pbits = 0, ubits = 0;
#pragma unroll
                            for (int b = U / kBatch - 1; b >= 0; b--) {
                                Slot pw[kBatch];
#pragma unroll
                                for (int u = 0; u < kBatch; u++)
                                    pw[u] = st_f[first + (b * kBatch + u) * G];
#pragma unroll
                                for (int u = kBatch - 1; u >= 0; u--) {
                                    float d2 = 0.0f;
#pragma unroll
                                    for (int k2 = 0; k2 < D; k2++) {
                                        float dk = ri[k2] - pw[u].v[k2];
                                        if (FOLD) dk = fmaf(-Lf[k2], rintf(dk * iLf[k2]), dk);
                                        d2 = fmaf(dk, dk, d2);
                                    }
                                    float c2lo = c2lou, c2hi = c2hiu;
                                    if constexpr (!UNI) {
                                        const float cc =
                                            fmaf(0.5f, dif + pw[u].v[UNI ? 0 : D], skf);
                                        const float c2 = cc * cc;
                                        c2lo = c2 - tol;
                                        c2hi = c2 + tol;
                                    }
                                    pbits = __builtin_amdgcn_alignbit(
                                        pbits, __float_as_uint(d2 - c2lo), 31u);
                                    ubits = __builtin_amdgcn_alignbit(
                                        ubits, __float_as_uint(d2 - c2hi), 31u);
                                }
                                if (b > 0) __asm__ volatile("" ::: "memory");
                            }
                            pbits &= valid;
                            ubits &= valid & ~pbits;
                            // candidates within the tolerance of the cutoff: the
                            // reference's exact predicate (one wave-uniform branch)
                            if (__ballot(ubits != 0u)) {
                                while (ubits) {
                                    const int u = __builtin_ctz(ubits);
                                    ubits &= ubits - 1u;
                                    const int s = first + u * G;
                                    const int cj = slot_cell(L, s);
                                    const int j = L.g0[cj] + (s - L.off[cj]);
                                    if (in_list<D>(xi, PosStore<D>::load(pos[j]), di, diam[j], box,
                                                   skin))
                                        pbits |= 1u << u;
                                }
                            }
                            // ignored pairs never enter the list (src/trackers.cpp:64)
                            if (ex.any) {
                                uint32_t b = pbits;
                                while (b) {
                                    const int u = __builtin_ctz(b);
                                    b &= b - 1u;
                                    const int s = first + u * G;
                                    const int cj = slot_cell(L, s);
                                    const int j = L.g0[cj] + (s - L.off[cj]);
                                    if (excluded(ex, oi, ex.orig[j])) pbits &= ~(1u << u);
                                }
                            }
                            uint32_t bits = pbits;
                            const int np = __builtin_popcount(bits);
                            // exclusive prefix of the round's passes over the group
                            // (the G lanes of one atom), and the group total
                            int incl, tot;
                            group_scan<G>(np, sub, glast, incl, tot);
                            int q = cnt + incl - np;
                            cnt += tot;
                            // entries at or past the capacity are not stored (the
                            // count is reported and the build repeated with more):
                            // only the lane's lowest cap - q passes stay
                            if (q + np > cap) {
                                uint32_t b = bits;
                                bits = 0u;
                                for (int k = max(cap - q, 0); k > 0 && b; k--) {
                                    const uint32_t low = b & (0u - b);
                                    bits |= low;
                                    b ^= low;
                                }
                            }
                            // the lane's passes in slot order, one store each: the
                            // wavefront runs as many store rounds as its busiest
                            // lane has passes (a few), not one per candidate
                            while (bits) {
                                const int u = __builtin_ctz(bits);
                                bits &= bits - 1u;
                                {
                                    const int s = first + u * G;
                                    if (off24) {
                                        const uint32_t qq = (uint32_t)q;
                                        uint32_t e;
                                        if (OUT32)
                                            e = __umul24(qq >> ts, rowe) + ie +
                                                (qq & (uint32_t)(tpa - 1));
                                        else
                                            e = __umul24(qq >> (ts + 1), rowe) + ie +
                                                ((qq & (uint32_t)(tpa - 1)) << 1) +
                                                ((qq >> ts) & 1u);
                                        if (OUT32)
                                            *reinterpret_cast<uint32_t *>(
                                                static_cast<char *>(nbr) + (e << 2)) = st_e[s];
                                        else
                                            *reinterpret_cast<uint16_t *>(
                                                static_cast<char *>(nbr) + (e << 1)) = (uint16_t)s;
                                    } else if (OUT32) {
                                        static_cast<uint32_t *>(
                                            nbr)[untiled_entry(q, i, tpa, npad)] = st_e[s];
                                    } else {
                                        static_cast<uint16_t *>(nbr)[tiled_entry(q, i, tpa, npad)] =
                                            (uint16_t)s;
                                    }
                                }
                                q++;
                            }
                        }
                    }
                }
            };
            if (anyfold)
                scan(std::true_type{});
            else
                scan(std::false_type{});
            if (sub == 0) {
                const int cc = cnt < cap ? cnt : cap;
                ncount[i] = cc;
                for (int l = 0; l < tpa; l++)
                    lcount[(size_t)i * tpa + l] =
                        (uint16_t)(cc > l ? (cc - l + tpa - 1) / tpa : 0);
                if (cnt > cap) atomicOr(flag, kFlagOverflow);
                mysum = (unsigned long long)cnt;
                myreq = (unsigned long long)cnt;
            }
        }
    }
    unsigned long long sum = mysum, mx = myreq;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off);
        const unsigned long long o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
    }
    if (lane == 0) {
        red[threadIdx.x / kWave][0] = sum;
        red[threadIdx.x / kWave][1] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s2 = 0, m2 = 0;
        for (int q = 0; q < G; q++) {
            s2 += red[q][0];
            m2 = red[q][1] > m2 ? red[q][1] : m2;
        }
        part[2 * (size_t)blockIdx.x] = s2;
        part[2 * (size_t)blockIdx.x + 1] = m2;
    }
}

// Tiled list -> untiled entries ((j << type_bits) | type_j at lane sub's
// entry m: (m * npad + i) * tpa + sub), for exporting the pairs.  One
// workgroup per tile.
template <int D>
__global__ __launch_bounds__(kBlock) void k_tiled_to_global(
    const TileRec *__restrict__ tiles, const uint32_t *__restrict__ keys,
    const int32_t *__restrict__ cstart, CellGrid g, const uint16_t *__restrict__ nb16,
    const uint16_t *__restrict__ lcount, const int32_t *__restrict__ type, int type_bits,
    int tpa, int npad, uint32_t *__restrict__ out) {
    __shared__ TileLayout L;
    const int t = blockIdx.x;
    tile_layout<D>(tiles, t, keys, cstart, g, L);
    if (L.row < 0) return;
    for (int e = threadIdx.x; e < L.cnt * tpa; e += kBlock) {
        const int i = L.a0 + e / tpa, sub = e % tpa;
        const int nl = lcount[(size_t)i * tpa + sub];
        for (int m = 0; m < nl; m++) {
            const int s = nb16[tiled_entry(m * tpa + sub, i, tpa, npad)];
            const int c = slot_cell(L, s);
            const int j = L.g0[c] + (s - L.off[c]);
            out[((size_t)m * npad + i) * tpa + sub] = ((uint32_t)j << type_bits) | (uint32_t)type[j];
        }
    }
}

// Wrapped copy of the positions, frac(x / L) * L per component, computed
// with the same floor as the cell keys: the list build's prefilter uses it.
template <int D>
__global__ __launch_bounds__(kBlock) void k_wrap(int n,
                                                 const typename PosStore<D>::T *__restrict__ pos,
                                                 CellGrid g, Box box,
                                                 typename PosStore<D>::T *__restrict__ xw) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Vec<D> x = PosStore<D>::load(pos[i]);
#pragma unroll
    for (int k = 0; k < D; k++) {
        double s = x.c[k] * g.invL[k];
        s -= floor(s);
        x.c[k] = s * box.L[k];
    }
    xw[i] = PosStore<D>::make(x);
}

constexpr int kStatsThreads = 1024;

__global__ __launch_bounds__(kStatsThreads) void k_build_stats(
    int nparts, const unsigned long long *__restrict__ part,
    unsigned long long *__restrict__ stats, uint32_t *__restrict__ flag,
    const int *__restrict__ info, BuildReport *report) {
    __shared__ unsigned long long red[kStatsThreads / kWave][2];
    unsigned long long sum = 0, mx = 0;
#pragma unroll 4
    for (int b = threadIdx.x; b < nparts; b += kStatsThreads) {
        sum += part[2 * (size_t)b];
        mx = part[2 * (size_t)b + 1] > mx ? part[2 * (size_t)b + 1] : mx;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off);
        const unsigned long long o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
    }
    const int lane = threadIdx.x % kWave;
    if (lane == 0) {
        red[threadIdx.x / kWave][0] = sum;
        red[threadIdx.x / kWave][1] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s2 = 0, m2 = 0;
        for (int q = 0; q < kStatsThreads / kWave; q++) {
            s2 += red[q][0];
            m2 = red[q][1] > m2 ? red[q][1] : m2;
        }
        stats[0] = s2;
        stats[1] = m2;
        const uint32_t fl = *flag;
        *flag = 0;
        if (report) {
            __hip_atomic_store(&report->sum, s2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&report->max, m2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&report->flag, fl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&report->ntiles, info ? info[0] : 0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&report->smax, info ? info[1] : 0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&report->bad, info ? info[2] : 0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// ======================================================== host launchers
static inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

hipError_t launch_cell_keys(int ndim, int n, const void *pos, const int32_t *rank,
                            const CellGrid &g, uint32_t *keys, int32_t *vals,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (ndim == 3)
        hipLaunchKernelGGL(k_cell_keys<3>, dim3(nblk(n)), dim3(kBlock), 0, s, n,
                           (const double4 *)pos, rank, g, keys, vals);
    else
        hipLaunchKernelGGL(k_cell_keys<2>, dim3(nblk(n)), dim3(kBlock), 0, s, n,
                           (const double2 *)pos, rank, g, keys, vals);
    return hipGetLastError();
}

int radix_blocks(int n) { return (n + kRadixTile - 1) / kRadixTile; }

// Sorts (keys, vals) by the low `bits` bits of the key.  The passes
// ping-pong between (keys, vals) and (tmp_keys, tmp_vals); after an odd
// number of passes the two pairs of pointers are swapped, so the result is
// in (keys, vals) either way.  hist holds 256 * (radix_blocks(n) + 1)
// counters (the last 256: digit totals).
hipError_t radix_sort(int n, int bits, uint32_t *&keys, int32_t *&vals,
                      uint32_t *&tmp_keys, int32_t *&tmp_vals, uint32_t *hist,
                      hipStream_t s) {
    if (n <= 1) return hipSuccess;
    const int nb = radix_blocks(n);
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    for (int p = 0; p < passes; p++) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kBlock), 0, s, n, shift,
                           keys, hist);
        hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(kBlock), 0, s, nb, hist,
                           hist + (size_t)256 * nb);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kBlock), 0, s, n, shift,
                           keys, vals, hist, hist + (size_t)256 * nb, tmp_keys, tmp_vals);
        std::swap(keys, tmp_keys);
        std::swap(vals, tmp_vals);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t gather(int n, const int32_t *perm, const T *src, T *dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather<T>, dim3(nblk(n)), dim3(kBlock), 0, s, n, perm, src,
                       dst);
    return hipGetLastError();
}
template hipError_t gather<double>(int, const int32_t *, const double *, double *,
                                   hipStream_t);
template hipError_t gather<int32_t>(int, const int32_t *, const int32_t *, int32_t *,
                                    hipStream_t);
template hipError_t gather<double2>(int, const int32_t *, const double2 *, double2 *,
                                    hipStream_t);
template hipError_t gather<double4>(int, const int32_t *, const double4 *, double4 *,
                                    hipStream_t);

hipError_t launch_permute(int ndim, int n, const int32_t *perm, const PermuteSet &P,
                          hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (ndim == 3)
        hipLaunchKernelGGL(k_permute<3>, dim3(nblk(n)), dim3(kBlock), 0, s, n, perm, P);
    else
        hipLaunchKernelGGL(k_permute<2>, dim3(nblk(n)), dim3(kBlock), 0, s, n, perm, P);
    return hipGetLastError();
}

hipError_t launch_cell_bounds(int n, int ncell, const uint32_t *keys, int32_t *start,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_cell_bounds, dim3(nblk(n + 1)), dim3(kBlock), 0, s, n,
                       ncell, keys, start);
    return hipGetLastError();
}

constexpr int kBuildLanes = 8;   // lanes per slot in k_build_list
constexpr int kBuildUnroll = 4;  // candidates per lane per round

int build_blocks(int n) { return (int)(((long)n * kBuildLanes + kBlock - 1) / kBlock); }

hipError_t launch_build_list(int ndim, int n, const void *pos, void *xw,
                             const int32_t *rank, const double *diam, const int32_t *type,
                             const uint32_t *keys, const int32_t *cstart,
                             const CellGrid &g, const Box &box, double skin, double dmax,
                             int type_bits, int tpa, int cap, int npad, uint32_t *nbr,
                             int32_t *ncount, uint16_t *lcount, void *xref,
                             unsigned long long *part, unsigned long long *stats,
                             uint32_t *flag, const ExclView &ex, BuildReport *report,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int nb = build_blocks(n);
    if (ndim == 3) {
        hipLaunchKernelGGL(k_wrap<3>, dim3(nblk(n)), dim3(kBlock), 0, s, n,
                           (const double4 *)pos, g, box, (double4 *)xw);
        hipLaunchKernelGGL((k_build_list<3, kBuildLanes, kBuildUnroll>), dim3(nb), dim3(kBlock), 0, s,
                           n, (const double4 *)pos, (const double4 *)xw, rank, diam, type,
                           keys, cstart, g, box, skin, dmax, type_bits, tpa, cap, npad,
                           nbr, ncount, lcount, (double4 *)xref, part, flag, ex);
    } else {
        hipLaunchKernelGGL(k_wrap<2>, dim3(nblk(n)), dim3(kBlock), 0, s, n,
                           (const double2 *)pos, g, box, (double2 *)xw);
        hipLaunchKernelGGL((k_build_list<2, kBuildLanes, kBuildUnroll>), dim3(nb), dim3(kBlock), 0, s,
                           n, (const double2 *)pos, (const double2 *)xw, rank, diam, type,
                           keys, cstart, g, box, skin, dmax, type_bits, tpa, cap, npad,
                           nbr, ncount, lcount, (double2 *)xref, part, flag, ex);
    }
    hipLaunchKernelGGL(k_build_stats, dim3(1), dim3(kStatsThreads), 0, s, nb, part, stats,
                       flag, (const int *)nullptr, report);
    return hipGetLastError();
}

// ------------------------------------------------------ tiled launchers
hipError_t launch_tiles(int ndim, int n, const CellGrid &g, const uint32_t *keys,
                        const int32_t *cstart, int cap_tiles, TileRec *tiles, int2 *truns,
                        int *info, hipStream_t s) {
    const int nrows = g.ncell / g.nc[0];
    hipLaunchKernelGGL(k_tile_rows, dim3(1), dim3(kTileScan), 0, s, n, nrows, g.nc[0],
                       g.ncell, cstart, cap_tiles, tiles, info);
    const int nb = (cap_tiles + kBlock - 1) / kBlock;
    if (ndim == 3)
        hipLaunchKernelGGL(k_tile_sizes<3>, dim3(nb), dim3(kBlock), 0, s, cap_tiles, g, keys,
                           cstart, tiles, truns, info);
    else
        hipLaunchKernelGGL(k_tile_sizes<2>, dim3(nb), dim3(kBlock), 0, s, cap_tiles, g, keys,
                           cstart, tiles, truns, info);
    return hipGetLastError();
}

constexpr int kTiledLanes = 4;   // lanes per atom in k_build_tiled
// candidates per lane per round: a stencil row (3 cells) of a dense system
// (the LJ liquid: ~50 atoms) fits one round of 4 x 16 (one prefix scan and
// one store loop for the row; 4 x 8 took two of each and was 7 % slower at
// 1M atoms); a sparse one (packed discs / spheres: ~7-9 atoms) one round of
// 4 x 2
constexpr int kTiledUnroll = 16, kTiledUnrollSparse = 2;

size_t tiled_build_lds(int smax) {
    return (size_t)(smax + kBuildPad) * sizeof(float4) + (size_t)smax * sizeof(uint32_t);
}

hipError_t launch_build_tiled(int ndim, int n, int ntiles, int smax, const TileRec *tiles,
                              const int2 *truns, const void *pos, void *xw, const int32_t *rank,
                              const double *diam, const int32_t *type, int type_bits,
                              const uint32_t *keys, const int32_t *cstart, const CellGrid &g,
                              const Box &box, double skin, double dmax, int tpa, int cap,
                              int npad, bool out32, void *nbr, int32_t *ncount,
                              uint16_t *lcount, void *xref, unsigned long long *part,
                              unsigned long long *stats, uint32_t *flag, const int *spec,
                              int uni, const ExclView &ex, BuildReport *report,
                              hipStream_t s) {
    if (n <= 0 || ntiles <= 0) return hipSuccess;
    const size_t slot = (size_t)(uni ? ndim : ndim + 1) * sizeof(float);
    const size_t lds = (size_t)(smax + kBuildPad) * slot +
                       (out32 ? (size_t)smax * sizeof(uint32_t) : 0);
    const bool sparse = 3.0 * (double)n / (double)g.ncell <= 12.0;  // atoms per stencil row
    // the build's 32-bit entry addressing (see k_build_tiled): the list under
    // 4 GiB and one entry row of npad atoms under 2^24 entries
    const int ts = 31 - __builtin_clz((unsigned)std::max(tpa, 1));
    const size_t esize = out32 ? 4 : 2;
    const size_t row = out32 ? ((size_t)npad << ts) : ((size_t)npad << (ts + 1));
    const int off24 = (size_t)cap * (size_t)npad * esize <= 0xffffffffull && row < (1u << 24);
#define PARM_BT(DD, PT, O32, UU, UNR)                                                        \
    hipLaunchKernelGGL((k_build_tiled<DD, kTiledLanes, UNR, O32, UU>), dim3(ntiles),          \
                       dim3(kWave * kTiledLanes), lds, s, tiles, truns, smax, (const PT *)pos, \
                       (const PT *)xw, rank, diam, type, type_bits, keys, cstart, g, box,     \
                       skin, dmax, tpa, cap, npad, off24, nbr, ncount, lcount, (PT *)xref,    \
                       part, flag, spec, ex)
#define PARM_BT_S(DD, PT, O32, UU)                       \
    if (sparse) {                                       \
        PARM_BT(DD, PT, O32, UU, kTiledUnrollSparse);   \
    } else {                                            \
        PARM_BT(DD, PT, O32, UU, kTiledUnroll);         \
    }
#define PARM_BT_U(DD, PT, O32)        \
    if (uni) {                       \
        PARM_BT_S(DD, PT, O32, true)  \
    } else {                         \
        PARM_BT_S(DD, PT, O32, false) \
    }
    if (ndim == 3) {
        hipLaunchKernelGGL(k_wrap<3>, dim3(nblk(n)), dim3(kBlock), 0, s, n,
                           (const double4 *)pos, g, box, (double4 *)xw);
        if (out32) {
            PARM_BT_U(3, double4, true)
        } else {
            PARM_BT_U(3, double4, false)
        }
    } else {
        hipLaunchKernelGGL(k_wrap<2>, dim3(nblk(n)), dim3(kBlock), 0, s, n,
                           (const double2 *)pos, g, box, (double2 *)xw);
        if (out32) {
            PARM_BT_U(2, double2, true)
        } else {
            PARM_BT_U(2, double2, false)
        }
    }
#undef PARM_BT_U
#undef PARM_BT_S
#undef PARM_BT
    hipLaunchKernelGGL(k_build_stats, dim3(1), dim3(kStatsThreads), 0, s, ntiles, part, stats,
                       flag, spec, report);
    return hipGetLastError();
}

hipError_t launch_tiled_to_global(int ndim, int ntiles, const TileRec *tiles,
                                  const uint32_t *keys, const int32_t *cstart,
                                  const CellGrid &g, const uint16_t *nb16,
                                  const uint16_t *lcount, const int32_t *type, int type_bits,
                                  int tpa, int npad, uint32_t *out, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    if (ndim == 3)
        hipLaunchKernelGGL(k_tiled_to_global<3>, dim3(ntiles), dim3(kBlock), 0, s, tiles,
                           keys, cstart, g, nb16, lcount, type, type_bits, tpa, npad, out);
    else
        hipLaunchKernelGGL(k_tiled_to_global<2>, dim3(ntiles), dim3(kBlock), 0, s, tiles,
                           keys, cstart, g, nb16, lcount, type, type_bits, tpa, npad, out);
    return hipGetLastError();
}

}  // namespace parm
