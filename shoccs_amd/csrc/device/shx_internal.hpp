// shx_internal.hpp -- device-side data layout and per-point operator evaluation
// shared by the HIP translation units of the SHOCCS device library (gfx950).
//
// Layout in HBM (mesh extents n0 x n1 x n2, z fastest: idx = (i*n1 + j)*n2 + k,
// as src/mesh/mesh.hpp:64-68):
//   * fields: D[N] then Rx, Ry, Rz, all FP64, one allocation per buffer;
//   * per direction, per line pair (slow, fast): a 64-bit descriptor holding the
//     two circulant row intervals of the line's first two segments
//     [lo0, hi0) and [lo1, hi1) (16 bits each).  A point whose position lies in
//     one of them takes the interior stencil with no lookup;
//   * per direction: pair_ptr[npairs + 1] and the segment table
//     (shx_line_block, 64 B each) for everything else: closure rows, B terms of
//     Dirichlet bodies, points with no row (solid points, removed Dirichlet rows,
//     lines lying on a Dirichlet face of another direction);
//   * a coefficient pool (closure rows and B weights), FP64.
//
// All arithmetic that reproduces a reference sum is written in the reference's
// order (dot products from 0.0, left to right; B terms added one at a time)
// and the library is compiled with -ffp-contract=off, so FP64 results are the
// same doubles as the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "shoccs_hip.h"

namespace shx {

constexpr int kTileZ = 64;  // lanes along the contiguous axis (one wavefront)
constexpr int kTileY = 4;   // wavefronts per workgroup, stacked along y
constexpr int kBlock = kTileZ * kTileY;
constexpr int kDerivPlanes = 4;  // x planes per lane of the point-wise derivative apply

// Device view of one direction's derivative operator (kernel argument).  Only
// scalars and pointers: a kernel argument that is indexed dynamically or whose
// address escapes into a call is copied to scratch memory for every thread.
struct DevDeriv {
    int dir;
    int nq;          // extent along dir
    int nfast;       // extent of the fast axis of the pair
    int width;       // interior stencil width
    long stride;     // D stride along dir
    const double* coef;  // interior stencil (device, `width` values)
    const int* pair_ptr;
    const unsigned long long* fast;
    const shx_line_block* blocks;
    const double* pool;
    int present;     // 0 when the direction has extent 1 (no operator)
};

// One row of a derivative at a point: the block dot product and up to two B
// products, in the order the reference adds them.
struct RowVal {
    double dot;
    double b0, b1;
    int nb;
    bool row;
};

__device__ __forceinline__ void point_coords(long idx, int n1, int n2, int& i, int& j, int& k)
{
    const long plane = (long)n1 * n2;
    i = (int)(idx / plane);
    const int rem = (int)(idx - (long)i * plane);
    j = rem / n2;
    k = rem - j * n2;
}

// (pair, q) of point (i, j, k) for direction `dir`, following
// index::dir<I>::slow/fast (src/indexing.hpp:21-37).
__device__ __forceinline__ void pair_of(int dir, int i, int j, int k, int n1, int n2, int& pair,
                                        int& q)
{
    if (dir == 0) {
        pair = j * n2 + k;
        q = i;
    } else if (dir == 1) {
        pair = i * n2 + k;
        q = j;
    } else {
        pair = i * n1 + j;
        q = k;
    }
}

// Row q of segment blk (0 <= q - blk.lo < its row count) at point idx: the
// closure or interior dot product from 0.0 in column order, then its B
// products.
template <bool WITH_B>
__device__ __forceinline__ RowVal row_in_block(const shx_line_block& blk, const double* __restrict__ pool,
                                               const double* __restrict__ coef, int width, long stride,
                                               const double* __restrict__ u,
                                               const double* __restrict__ uR, long idx, int q)
{
    RowVal r;
    r.b0 = r.b1 = 0.0;
    r.nb = 0;
    r.row = true;
    const int lr = q - blk.lo;
    const int total = blk.nl + blk.ni + blk.nr;
    const long base = idx - (long)q * stride;
    double dot = 0.0;
    if (lr < blk.nl) {
        const double* a = pool + blk.lcoef + (long)lr * blk.lcols;
        for (int jj = 0; jj < blk.lcols; ++jj)
            dot += a[jj] * u[base + (long)(blk.lcol0 + jj) * stride];
    } else if (lr < blk.nl + blk.ni) {
        const int hw = width / 2;
        for (int jj = 0; jj < width; ++jj) dot += coef[jj] * u[idx + (long)(jj - hw) * stride];
    } else {
        const int rr = lr - blk.nl - blk.ni;
        const double* a = pool + blk.rcoef + (long)rr * blk.rcols;
        for (int jj = 0; jj < blk.rcols; ++jj)
            dot += a[jj] * u[base + (long)(blk.rcol0 + jj) * stride];
    }
    r.dot = dot;
    if (WITH_B) {
        if (blk.lobj >= 0 && lr < blk.nl) {
            r.b0 = pool[blk.lbw + lr] * uR[blk.lobj];
            r.nb = 1;
        }
        if (blk.robj >= 0 && lr >= total - blk.nr) {
            const double v = pool[blk.rbw + (lr - (total - blk.nr))] * uR[blk.robj];
            if (r.nb == 0) r.b0 = v;
            else r.b1 = v;
            r.nb++;
        }
    }
    return r;
}

// Slow path: scan the segments of the line for the one holding row q.  All
// arguments are scalars or pointers so that nothing is spilled to scratch.
template <bool WITH_B>
__device__ __noinline__ RowVal row_slow(const int* __restrict__ pair_ptr,
                                        const shx_line_block* __restrict__ blocks,
                                        const double* __restrict__ pool,
                                        const double* __restrict__ coef, int width, long stride,
                                        const double* __restrict__ u,
                                        const double* __restrict__ uR, long idx, int pair, int q)
{
    const int b_begin = pair_ptr[pair];
    const int b_end = pair_ptr[pair + 1];
    for (int b = b_begin; b < b_end; ++b) {
        const shx_line_block blk = blocks[b];
        const int lr = q - blk.lo;
        if (lr < 0) break;
        if (lr >= blk.nl + blk.ni + blk.nr) continue;
        return row_in_block<WITH_B>(blk, pool, coef, width, stride, u, uR, idx, q);
    }
    RowVal r;
    r.dot = 0.0;
    r.b0 = r.b1 = 0.0;
    r.nb = 0;
    r.row = false;
    return r;
}

// A row of a derivative outside the fast intervals of its line: point idx,
// position q along the direction, segment b.  The point-wise apply evaluates
// these from a list (sorted by idx) in a launch of their own.
struct SlowRow {
    long idx;
    int b;
    int q;
};

// The closure rows of one end of a segment with no B term at that end (the
// grid walls, above all), evaluated by one lane: the end's columns are loaded
// once and every row of the end sums over them, from 0.0 in column order.
constexpr int kWallRows = 8, kWallCols = 8;
struct WallSide {
    long base;  // point of position 0 of the line
    int coef;   // pool offset of the rows (row-major, `cols` values per row)
    int q0;     // position of the first row
    int col0;   // position of the first column
    short rows, cols;
};

// Row of direction DIR (stencil width W) at point idx = (i, j, k).  The W taps
// are loaded before the interval test, at positions clamped to the line, so
// that their loads are issued together; they are used only when q lies in one
// of the line's interior intervals (where no clamping occurs).
template <int DIR, int W, bool WITH_B>
__device__ __forceinline__ RowVal row_at(const DevDeriv& d, const double* __restrict__ u,
                                         const double* __restrict__ uR, long idx, int i, int j,
                                         int k, int n1, int n2)
{
    constexpr int hw = W / 2;
    const int pair = DIR == 0 ? j * n2 + k : DIR == 1 ? i * n2 + k : i * n1 + j;
    const int q = DIR == 0 ? i : DIR == 1 ? j : k;
    const unsigned long long fd = d.fast[pair];
    double v[W];
#pragma unroll
    for (int jj = 0; jj < W; ++jj) {
        int qq = q + jj - hw;
        qq = qq < 0 ? 0 : (qq > d.nq - 1 ? d.nq - 1 : qq);
        v[jj] = u[idx + (long)(qq - q) * d.stride];
    }
    const int lo0 = (int)(fd & 0xffffu), hi0 = (int)((fd >> 16) & 0xffffu);
    const int lo1 = (int)((fd >> 32) & 0xffffu), hi1 = (int)(fd >> 48);
    if ((q >= lo0 && q < hi0) || (q >= lo1 && q < hi1)) {
        RowVal r;
        double dot = 0.0;
#pragma unroll
        for (int jj = 0; jj < W; ++jj) dot += d.coef[jj] * v[jj];
        r.dot = dot;
        r.b0 = r.b1 = 0.0;
        r.nb = 0;
        r.row = true;
        return r;
    }
    return row_slow<WITH_B>(d.pair_ptr, d.blocks, d.pool, d.coef, W, d.stride, u, uR, idx, pair, q);
}

// eq_t result of a derivative row: dot, then the B terms.
__device__ __forceinline__ double row_eq(const RowVal& r)
{
    double v = r.dot;
    if (r.nb > 0) v += r.b0;
    if (r.nb > 1) v += r.b1;
    return v;
}

// plus_eq_t: acc += dot, then the B terms.
__device__ __forceinline__ double row_plus(double acc, const RowVal& r)
{
    acc += r.dot;
    if (r.nb > 0) acc += r.b0;
    if (r.nb > 1) acc += r.b1;
    return acc;
}

// Workgroup -> tile mapping, contiguous per XCD: workgroup b runs on XCD b % 8;
// give every XCD a contiguous range of logical tiles so that the x-neighbour
// planes a tile reads are fetched into the same L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks)
{
    const unsigned per = nblocks / 8u, rem = nblocks % 8u;
    const unsigned xcd = b % 8u, local = b / 8u;
    return xcd < rem ? xcd * (per + 1u) + local : rem * (per + 1u) + (xcd - rem) * per + local;
}

struct Tiling {
    int n0, n1, n2;
    int nbz, nby;
    unsigned nblocks;
};

inline Tiling make_tiling(const int n[3])
{
    Tiling t;
    t.n0 = n[0];
    t.n1 = n[1];
    t.n2 = n[2];
    t.nbz = (n[2] + kTileZ - 1) / kTileZ;
    t.nby = (n[1] + kTileY - 1) / kTileY;
    t.nblocks = (unsigned)t.nbz * (unsigned)t.nby * (unsigned)n[0];
    return t;
}

// Point handled by this thread of tile workgroup b; returns false outside the
// mesh.
__device__ __forceinline__ bool tile_point(const Tiling& t, unsigned b, int& i, int& j, int& k, long& idx)
{
    const unsigned L = xcd_remap(b, t.nblocks);
    i = (int)(L % (unsigned)t.n0);
    const unsigned rest = L / (unsigned)t.n0;
    const int bz = (int)(rest % (unsigned)t.nbz);
    const int by = (int)(rest / (unsigned)t.nbz);
    k = bz * kTileZ + (int)(threadIdx.x % kTileZ);
    j = by * kTileY + (int)(threadIdx.x / kTileZ);
    if (k >= t.n2 || j >= t.n1) return false;
    idx = ((long)i * t.n1 + j) * t.n2 + k;
    return true;
}

// The same tile with each wave on all kTileY rows of a patch of
// 64 / kTileY points along z, for rows along z: the closure rows at the ends
// of a z line then fall in the end waves of the end tiles only, instead of
// diverging every wave of those tiles.
__device__ __forceinline__ bool tile_point_zpatch(const Tiling& t, unsigned b, int& i, int& j, int& k,
                                                  long& idx)
{
    constexpr int ZW = 64 / kTileY;
    static_assert(kTileZ == 64 && 64 % kTileY == 0, "a wave holds whole rows of its z patch");
    const unsigned L = xcd_remap(b, t.nblocks);
    i = (int)(L % (unsigned)t.n0);
    const unsigned rest = L / (unsigned)t.n0;
    const int bz = (int)(rest % (unsigned)t.nbz);
    const int by = (int)(rest / (unsigned)t.nbz);
    k = bz * kTileZ + (int)(threadIdx.x / 64) * ZW + (int)(threadIdx.x % ZW);
    j = by * kTileY + (int)((threadIdx.x % 64) / ZW);
    if (k >= t.n2 || j >= t.n1) return false;
    idx = ((long)i * t.n1 + j) * t.n2 + k;
    return true;
}

// The same for the workgroups of a launch that holds tiles only.
__device__ __forceinline__ bool tile_point(const Tiling& t, int& i, int& j, int& k, long& idx)
{
    return tile_point(t, blockIdx.x, i, j, k, idx);
}

__device__ __forceinline__ bool tile_point_zpatch(const Tiling& t, int& i, int& j, int& k, long& idx)
{
    return tile_point_zpatch(t, blockIdx.x, i, j, k, idx);
}

}  // namespace shx
