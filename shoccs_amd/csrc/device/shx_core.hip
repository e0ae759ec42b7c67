// shx_core.hip -- context, memory, derivative / block / CSR operators,
// gradient and laplacian of the SHOCCS device library (gfx950).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "shx_internal.hpp"
#include "shx_state.hpp"

using namespace shx;

// ------------------------------------------------------------------ kernels
namespace {

__global__ void k_fill(double* __restrict__ dst, double v, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

// two doubles per lane: 16-byte loads and stores, grid-stride over the pairs
__global__ void k_copy_vec(double2* __restrict__ dst, const double2* __restrict__ src, long npairs)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npairs;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_copy_tail(double* __restrict__ dst, const double* __restrict__ src, long n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && (n & 1)) dst[n - 1] = src[n - 1];
}

// The rows outside the fast intervals, evaluated by the first nlb
// workgroups of the interior apply's launch: lane i takes listed row i
// (closures with B terms, interior rows of a line's third and later
// segments) for i < n; the L = 2^lsh lanes n + L w .. n + L w + L - 1 take
// closure end w, lane r of them its rows r, r + L, ...  Along x and y one
// lane takes an end (neighbouring lanes take neighbouring lines, whose
// columns are adjacent in memory); along z four (the four load the same
// columns, so that a load instruction touches 16 lines, not 64).
struct ListArgs {
    long n;
    const SlowRow* rows;
    long nw;
    const WallSide* walls;
    int lsh;
    unsigned nlb;  // workgroups of the list
};

template <int W, bool PLUS, bool WITH_B>
__device__ __forceinline__ void list_item(long i, const ListArgs& la, const DevDeriv& d,
                                          const double* __restrict__ u, const double* __restrict__ uR,
                                          double* __restrict__ du)
{
    if (i < la.n) {
        const SlowRow e = la.rows[i];
        const RowVal r = row_in_block<WITH_B>(d.blocks[e.b], d.pool, d.coef, W, d.stride, u, uR, e.idx, e.q);
        du[e.idx] = PLUS ? row_plus(du[e.idx], r) : row_eq(r);
        return;
    }
    const long w = (i - la.n) >> la.lsh;
    const int r0 = (int)((i - la.n) & ((1 << la.lsh) - 1));
    if (w >= la.nw) return;
    const WallSide e = la.walls[w];
    const long st = d.stride;
    double v[kWallCols];
#pragma unroll
    for (int m = 0; m < kWallCols; ++m) v[m] = m < e.cols ? u[e.base + (long)(e.col0 + m) * st] : 0.0;
    for (int r = r0; r < kWallRows; r += 1 << la.lsh) {
        if (r < e.rows) {
            const double* a = d.pool + e.coef + r * e.cols;
            double dot = 0.0;
#pragma unroll
            for (int m = 0; m < kWallCols; ++m)
                if (m < e.cols) dot += a[m] * v[m];
            const long idx = e.base + (long)(e.q0 + r) * st;
            du[idx] = PLUS ? du[idx] + dot : dot;
        }
    }
}

// Interior rows of the fast intervals.  A lane evaluates the points of its
// (y, z) position in XP consecutive x planes.  Every point's taps (and, for
// plus_eq, its old value) are loaded before any row is evaluated, so that a
// lane keeps several points' loads in flight; rows along x share one window
// of planes.  Each row is summed as row_at sums it.  With XP > 1 (large
// grids) every other row comes from the slow-row list and the closure ends
// (list_item, in workgroups after the tiles), so no wave waits on a segment
// scan; with XP = 1 (grids too
// small to fill the chip, where a second launch costs more than the
// divergence) the lane scans its line's segments itself.  Points with no row
// are left as they are.  t counts groups of XP planes along x.
template <int DIR, int W, bool PLUS, bool WITH_B, int XP>
__global__ __launch_bounds__(kBlock) void k_derivative(Tiling t, int n0, DevDeriv d,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ uR,
                                                       double* __restrict__ du, ListArgs la)
{
    constexpr int hw = W / 2;
    unsigned bid = blockIdx.x;
    if constexpr (XP > 1) {
        if (bid < la.nlb) {
            list_item<W, PLUS, WITH_B>((long)bid * kBlock + threadIdx.x, la, d, u, uR, du);
            return;
        }
        bid -= la.nlb;
    }
    int g, j, k;
    long unused;
    if (!(DIR == 2 ? tile_point_zpatch(t, bid, g, j, k, unused) : tile_point(t, bid, g, j, k, unused))) return;
    if constexpr (DIR == 1 && XP > 1) {
        // rows along y: the tiling counts groups of XP rows (t.n1 groups), a
        // lane takes the XP consecutive rows of group j in its (x, z) column,
        // and all their taps come from one window of rows (clamped to the
        // line; the clamped values only reach rows outside the fast
        // intervals, which list_item evaluates)
        const int ny = d.nq, nz = t.n2;
        const int j0 = j * XP;
        const long colb = (long)g * ny * nz + k;  // (g, 0, k)
        const unsigned long long f = d.fast[g * nz + k];
        double win[XP + W - 1];
#pragma unroll
        for (int m = 0; m < XP + W - 1; ++m) win[m] = u[colb + (long)min(max(j0 + m - hw, 0), ny - 1) * nz];
        double old[XP];
#pragma unroll
        for (int p = 0; p < XP; ++p) old[p] = PLUS ? du[colb + (long)min(j0 + p, ny - 1) * nz] : 0.0;
        const int lo0 = (int)(f & 0xffffu), hi0 = (int)((f >> 16) & 0xffffu);
        const int lo1 = (int)((f >> 32) & 0xffffu), hi1 = (int)(f >> 48);
#pragma unroll
        for (int p = 0; p < XP; ++p) {
            const int q = j0 + p;
            if (q >= ny) break;
            if (!((q >= lo0 && q < hi0) || (q >= lo1 && q < hi1))) continue;
            double dot = 0.0;
#pragma unroll
            for (int m = 0; m < W; ++m) dot += d.coef[m] * win[p + m];
            du[colb + (long)q * nz] = PLUS ? old[p] + dot : dot;
        }
        return;
    }
    const int n1 = t.n1, n2 = t.n2;
    const long plane = (long)n1 * n2, col = (long)j * n2 + k;
    const int i0 = g * XP;
    long idx[XP];
    int pair[XP];
    double v[XP][W];
#pragma unroll
    for (int p = 0; p < XP; ++p) {
        const int ic = min(i0 + p, n0 - 1);
        idx[p] = (long)ic * plane + col;
        pair[p] = DIR == 0 ? j * n2 + k : DIR == 1 ? ic * n2 + k : ic * n1 + j;
    }
    unsigned long long fd[XP];
    if constexpr (DIR == 0) {
        const unsigned long long f = d.fast[pair[0]];
        double win[XP + W - 1];
#pragma unroll
        for (int m = 0; m < XP + W - 1; ++m) {
            const int qq = min(max(i0 + m - hw, 0), d.nq - 1);
            win[m] = u[(long)qq * plane + col];
        }
#pragma unroll
        for (int p = 0; p < XP; ++p) {
            fd[p] = f;
#pragma unroll
            for (int m = 0; m < W; ++m) v[p][m] = win[p + m];
        }
    } else {
        const int q = DIR == 1 ? j : k;
#pragma unroll
        for (int p = 0; p < XP; ++p) {
            fd[p] = d.fast[pair[p]];
#pragma unroll
            for (int m = 0; m < W; ++m) {
                const int qq = min(max(q + m - hw, 0), d.nq - 1);
                v[p][m] = u[idx[p] + (long)(qq - q) * d.stride];
            }
        }
    }
    double old[XP];
#pragma unroll
    for (int p = 0; p < XP; ++p) old[p] = PLUS ? du[idx[p]] : 0.0;
#pragma unroll
    for (int p = 0; p < XP; ++p) {
        if (i0 + p >= n0) break;
        const int q = DIR == 0 ? i0 + p : DIR == 1 ? j : k;
        const int lo0 = (int)(fd[p] & 0xffffu), hi0 = (int)((fd[p] >> 16) & 0xffffu);
        const int lo1 = (int)((fd[p] >> 32) & 0xffffu), hi1 = (int)(fd[p] >> 48);
        if (!((q >= lo0 && q < hi0) || (q >= lo1 && q < hi1))) {
            if constexpr (XP == 1) {
                const RowVal r = row_slow<WITH_B>(d.pair_ptr, d.blocks, d.pool, d.coef, W, d.stride, u, uR,
                                                  idx[p], pair[p], q);
                if (r.row) du[idx[p]] = PLUS ? row_plus(old[p], r) : row_eq(r);
            }
            continue;
        }
        double dot = 0.0;
#pragma unroll
        for (int m = 0; m < W; ++m) dot += d.coef[m] * v[p][m];
        du[idx[p]] = PLUS ? old[p] + dot : dot;
    }
}

// VEC doubles as one vector (aligned loads and stores)
template <int VEC>
struct VecT {
    typedef double type __attribute__((ext_vector_type(VEC)));
};

// Interior rows along z, the contiguous axis, when n2 is a multiple of VEC: a
// lane takes VEC consecutive points of NL lines (i, j) and a wave 64 VEC
// points of a line.  The taps come from three aligned VEC-vectors per line
// (the lane's own and its neighbours' positions, clamped into the line; the
// clamped values only reach rows outside the fast intervals), so a point
// costs 3 / VEC loads instead of W.  Each row is summed as row_at sums it;
// every row outside the line's fast intervals is left to list_item, in the
// workgroups after the first la.first.
template <int W, bool PLUS, bool WITH_B, int VEC, int NL>
__global__ __launch_bounds__(256) void k_derivative_z(long nlines, int n2, int nzc, DevDeriv d,
                                                      const double* __restrict__ u,
                                                      const double* __restrict__ uR, double* __restrict__ du,
                                                      ListArgs la)
{
    constexpr int hw = W / 2;
    static_assert(hw <= VEC, "the taps lie in the three vectors");
    using V = typename VecT<VEC>::type;
    if (blockIdx.x < la.nlb) {
        list_item<W, PLUS, WITH_B>((long)blockIdx.x * 256 + threadIdx.x, la, d, u, uR, du);
        return;
    }
    const unsigned b = xcd_remap(blockIdx.x - la.nlb, gridDim.x - la.nlb);
    const long task = (long)b * 4 + threadIdx.x / 64;  // one wave: one z chunk of NL lines
    const int zc = (int)(task % nzc);
    const long l0 = (task / nzc) * NL;
    const int k0 = zc * 64 * VEC + (int)(threadIdx.x % 64) * VEC;
    if (k0 >= n2 || l0 >= nlines) return;
    const int kl = max(k0 - VEC, 0), kr = min(k0 + VEC, n2 - VEC);
    double v[NL][3 * VEC], old[NL][VEC];
    unsigned long long fd[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const long base = min(l0 + l, nlines - 1) * n2;
        const V a = *reinterpret_cast<const V*>(u + base + kl);
        const V c = *reinterpret_cast<const V*>(u + base + k0);
        const V r = *reinterpret_cast<const V*>(u + base + kr);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            v[l][e] = a[e];
            v[l][VEC + e] = c[e];
            v[l][2 * VEC + e] = r[e];
        }
        fd[l] = d.fast[min(l0 + l, nlines - 1)];
        if (PLUS) {
            const V o = *reinterpret_cast<const V*>(du + base + k0);
#pragma unroll
            for (int e = 0; e < VEC; ++e) old[l][e] = o[e];
        }
    }
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        if (l0 + l >= nlines) break;
        const int lo0 = (int)(fd[l] & 0xffffu), hi0 = (int)((fd[l] >> 16) & 0xffffu);
        const int lo1 = (int)((fd[l] >> 32) & 0xffffu), hi1 = (int)(fd[l] >> 48);
        V o;
        bool in[VEC];
        bool all = true;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            const int q = k0 + e;
            in[e] = (q >= lo0 && q < hi0) || (q >= lo1 && q < hi1);
            all = all && in[e];
            double dot = 0.0;
#pragma unroll
            for (int m = 0; m < W; ++m) dot += d.coef[m] * v[l][VEC + e - hw + m];
            o[e] = PLUS ? old[l][e] + dot : dot;
        }
        double* p = du + (l0 + l) * n2 + k0;
        if (all) {
            *reinterpret_cast<V*>(p) = o;
        } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e)
                if (in[e]) p[e] = o[e];
        }
    }
}

// N matrix rows, compressed to the rows that hold entries.
__global__ void k_rows_csr(long nrows, const long* __restrict__ rows, const int* __restrict__ ptr,
                           const long* __restrict__ col, const double* __restrict__ w,
                           const double* __restrict__ x, double* __restrict__ b)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const long row = rows[r];
    double acc = b[row];
    for (int e = ptr[r]; e < ptr[r + 1]; ++e) acc += w[e] * x[col[e]];
    b[row] = acc;
}

// R-space rows of non-Dirichlet bodies (the Bf and Br products of
// src/operators/derivative.cpp:479-485, always +=): one lane per row,
// du.R[set][row] += the row's entries over u.D (column >= 0) and u.R[set]
// (column -1 - c), summed in stored order.
__global__ void k_cut_rows(long nrows, const int* __restrict__ rset, const long* __restrict__ rrow,
                           const long* __restrict__ rptr, const long* __restrict__ rcol,
                           const double* __restrict__ rw, const double* __restrict__ uD,
                           const double* uR0, const double* uR1, const double* uR2, double* duR0,
                           double* duR1, double* duR2)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int s = rset[k];
    const double* uR = s == 0 ? uR0 : s == 1 ? uR1 : uR2;
    double* duR = s == 0 ? duR0 : s == 1 ? duR1 : duR2;
    const long row = rrow[k];
    double acc = duR[row];
    for (long e = rptr[k]; e < rptr[k + 1]; ++e) {
        const long c = rcol[e];
        acc += rw[e] * (c >= 0 ? uD[c] : uR[-1 - c]);
    }
    duR[row] = acc;
}

// The reference's csr apply over every row pointer (src/matrices/csr.cpp:37-51).
__global__ void k_csr_full(long nrows, const long* __restrict__ u, const long* __restrict__ v,
                           const double* __restrict__ w, const double* __restrict__ x,
                           double* __restrict__ b)
{
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const long e0 = u[row], e1 = u[row + 1];
    if (e0 == e1) return;
    double acc = b[row];
    for (long e = e0; e < e1; ++e) acc += w[e] * x[v[e]];
    b[row] = acc;
}

}  // namespace

// ------------------------------------------------------------------ helpers
namespace shx {

int fail(shx_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(shx_ctx* ctx, hipError_t e, const char* what)
{
    return fail(ctx, SHX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int launch_fill(shx_ctx* ctx, double* dst, double v, long n)
{
    if (n <= 0) return SHX_OK;
    const unsigned grid = (unsigned)((n + 255) / 256);
    k_fill<<<grid, 256, 0, ctx->stream>>>(dst, v, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "fill");
}

static int launch_copy(shx_ctx* ctx, double* dst, const double* src, long n)
{
    if (n <= 0) return SHX_OK;
    const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15) == 0;
    if (!aligned) return fail(ctx, SHX_EINVAL, "copy: buffers must be 16-byte aligned");
    const long pairs = n / 2;
    if (pairs > 0) {
        const long want = (pairs + 255) / 256;
        const unsigned grid = (unsigned)(want < 65536 ? want : 65536);
        k_copy_vec<<<grid, 256, 0, ctx->stream>>>(reinterpret_cast<double2*>(dst),
                                                  reinterpret_cast<const double2*>(src), pairs);
    }
    if (n & 1) k_copy_tail<<<1, 64, 0, ctx->stream>>>(dst, src, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "copy");
}

// The list's workgroups, appended to an interior apply of `first` workgroups.
ListArgs list_args(const shx_derivative* d, unsigned first, unsigned& grid)
{
    const int lsh = d->dir == 2 ? 2 : 0;
    const long lanes = d->nslow_list + (d->nwall << lsh);
    const ListArgs la{d->nslow_list, d->slow, d->nwall, d->walls, lsh, (unsigned)((lanes + 255) / 256)};
    grid = first + la.nlb;
    return la;
}

int launch_derivative(shx_derivative* d, const double* u, const double* uR, double* du, bool plus,
                      bool with_b)
{
    shx_ctx* ctx = d->ctx;
    if (!d->dev.present) return SHX_OK;
    // four planes (rows along y: four rows) per lane and the slow rows from
    // their list (measured at 256^3: x 4.3, y 4.3 TB/s, against 3.0, 3.0
    // with one plane and the segment scan in every lane), unless the grid
    // would then hold fewer than eight workgroups per CU
    const int g4[3] = {(d->n[0] + kDerivPlanes - 1) / kDerivPlanes, d->n[1], d->n[2]};
    const Tiling t4 = make_tiling(g4);
    const bool multi = t4.nblocks >= 8u * 256u;
    // rows along y take groups of kDerivPlanes consecutive rows per lane
    // instead of planes (one window of rows per lane)
    const int g4y[3] = {d->n[0], (d->n[1] + kDerivPlanes - 1) / kDerivPlanes, d->n[2]};
    const Tiling t = multi ? (d->dir == 1 ? make_tiling(g4y) : t4) : make_tiling(d->n);
    // rows along z: the vector form when the z extent allows aligned vectors
    // and the grid is large enough for a second launch to pay (at 256^3 it
    // takes the z apply from 2.5 to 4.0 TB/s of 16 B per point)
    if (d->dir == 2 && multi && d->n[2] % 2 == 0) {
        const int vec = d->n[2] % 4 == 0 ? 4 : 2;
        constexpr int NL = 2;
        const long nlines = (long)d->n[0] * d->n[1];
        const int nzc = (d->n[2] + 64 * vec - 1) / (64 * vec);
        const long waves = (nlines + NL - 1) / NL * nzc;
        unsigned grid = 0;
        const ListArgs la = list_args(d, (unsigned)((waves + 3) / 4), grid);
        const int zkey = d->dev.width * 1000 + vec * 100 + (plus ? 10 : 0) + (with_b ? 1 : 0);
        switch (zkey) {
#define SHX_DZ_CASE(W, V, P, B) \
    case W * 1000 + V * 100 + (P ? 10 : 0) + (B ? 1 : 0): \
        k_derivative_z<W, P, B, V, NL><<<grid, 256, 0, ctx->stream>>>(nlines, d->n[2], nzc, d->dev, u, uR, du, la); \
        break;
#define SHX_DZ_W(W) \
    SHX_DZ_CASE(W, 2, false, false) SHX_DZ_CASE(W, 2, false, true) SHX_DZ_CASE(W, 2, true, false) \
    SHX_DZ_CASE(W, 2, true, true) SHX_DZ_CASE(W, 4, false, false) SHX_DZ_CASE(W, 4, false, true) \
    SHX_DZ_CASE(W, 4, true, false) SHX_DZ_CASE(W, 4, true, true)
            SHX_DZ_W(3) SHX_DZ_W(5)
#undef SHX_DZ_W
#undef SHX_DZ_CASE
        default:
            return fail(ctx, SHX_EUNSUPPORTED, "derivative: interior width must be 3 or 5");
        }
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "derivative kernel");
    }
    static_assert(kBlock == 256, "the list's workgroups have the tiles' size");
    unsigned grid = 0;
    const ListArgs la = multi ? list_args(d, t.nblocks, grid) : ListArgs{0, nullptr, 0, nullptr, 0, 0};
    const int key = d->dir * 100 + d->dev.width * 10 + (plus ? 2 : 0) + (with_b ? 1 : 0);
    switch (key) {
#define SHX_DERIV_CASE(D, W, P, B) \
    case D * 100 + W * 10 + (P ? 2 : 0) + (B ? 1 : 0): \
        if (multi) \
            k_derivative<D, W, P, B, kDerivPlanes><<<grid, kBlock, 0, ctx->stream>>>(t, d->n[0], d->dev, u, uR, du, la); \
        else \
            k_derivative<D, W, P, B, 1><<<t.nblocks, kBlock, 0, ctx->stream>>>(t, d->n[0], d->dev, u, uR, du, la); \
        break;
#define SHX_DERIV_DIR(D, W) \
    SHX_DERIV_CASE(D, W, false, false) SHX_DERIV_CASE(D, W, false, true) \
    SHX_DERIV_CASE(D, W, true, false) SHX_DERIV_CASE(D, W, true, true)
        SHX_DERIV_DIR(0, 3) SHX_DERIV_DIR(1, 3) SHX_DERIV_DIR(2, 3)
        SHX_DERIV_DIR(0, 5) SHX_DERIV_DIR(1, 5) SHX_DERIV_DIR(2, 5)
#undef SHX_DERIV_DIR
#undef SHX_DERIV_CASE
    default:
        return fail(ctx, SHX_EUNSUPPORTED, "derivative: interior width must be 3 or 5");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "derivative kernel");
}

int launch_neumann(shx_derivative* d, const double* nu, double* du)
{
    if (d->nneu_rows == 0) return SHX_OK;
    shx_ctx* ctx = d->ctx;
    const unsigned grid = (unsigned)((d->nneu_rows + 255) / 256);
    k_rows_csr<<<grid, 256, 0, ctx->stream>>>(d->nneu_rows, d->neu_rows, d->neu_ptr, d->neu_col,
                                             d->neu_w, nu, du);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "neumann kernel");
}

int launch_cut_rows(shx_derivative* d, const shx_field* u, const shx_field* du)
{
    if (d->nrrow == 0) return SHX_OK;
    shx_ctx* ctx = d->ctx;
    const unsigned grid = (unsigned)((d->nrrow + 255) / 256);
    k_cut_rows<<<grid, 256, 0, ctx->stream>>>(d->nrrow, d->rset, d->rrow, d->rptr, d->rcol, d->rw, u->D,
                                             u->R[0], u->R[1], u->R[2], du->R[0], du->R[1], du->R[2]);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "cut-row kernel");
}

int launch_csr_full(shx_csr* A, const double* x, double* b)
{
    if (A->rows == 0) return SHX_OK;
    const unsigned grid = (unsigned)((A->rows + 255) / 256);
    k_csr_full<<<grid, 256, 0, A->ctx->stream>>>(A->rows, A->u, A->v, A->w, x, b);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(A->ctx, e, "csr kernel");
}

int zero_field(shx_ctx* ctx, const shx_field* f, const int n[3], const long nR[3])
{
    int rc = launch_fill(ctx, f->D, 0.0, (long)n[0] * n[1] * n[2]);
    for (int r = 0; r < 3 && rc == SHX_OK; ++r)
        if (nR[r] > 0) rc = launch_fill(ctx, f->R[r], 0.0, nR[r]);
    return rc;
}

int sync(shx_ctx* ctx)
{
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "stream synchronize");
}

template <typename T>
int upload(shx_ctx* ctx, T** dst, const T* src, size_t n)
{
    *dst = nullptr;
    if (n == 0) return SHX_OK;
    hipError_t e = hipMalloc((void**)dst, n * sizeof(T));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    e = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy");
    return SHX_OK;
}

template int upload<double>(shx_ctx*, double**, const double*, size_t);
template int upload<int>(shx_ctx*, int**, const int*, size_t);
template int upload<long>(shx_ctx*, long**, const long*, size_t);
template int upload<unsigned>(shx_ctx*, unsigned**, const unsigned*, size_t);
template int upload<int8_t>(shx_ctx*, int8_t**, const int8_t*, size_t);

}  // namespace shx

// ------------------------------------------------------------- workspace
// The workspace of the allocation-free entry points: one copy in the memory of
// every GPU the library runs on, held until the process ends.  In host code
// this name is only a stand-in for the symbol, never a device address; each
// device's address comes from hipGetSymbolAddress at shx_init (with that
// device current) and is kept per device below and in the context.
static __device__ __align__(256) unsigned char scratch[shx::kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_addr[kMaxDevices] = {};

// the current device's workspace address (resolved once per device)
hipError_t resolve_scratch(int device, void** addr)
{
    *addr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipSuccess;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = p;
    }
    *addr = g_scratch_addr[device];
    return hipSuccess;
}
}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int shx_init(int device, shx_ctx** out)
{
    if (!out) return SHX_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return SHX_EHIP;
    shx_ctx* ctx = new (std::nothrow) shx_ctx();
    if (!ctx) return SHX_ENOMEM;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = resolve_scratch(device, &ctx->scratch);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return SHX_EHIP;
    }
    *out = ctx;
    return SHX_OK;
}

int shx_finalize(shx_ctx* ctx)
{
    if (!ctx) return SHX_EINVAL;
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return SHX_OK;
}

const char* shx_last_error(const shx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int shx_sync(shx_ctx* ctx) { return ctx ? sync(ctx) : SHX_EINVAL; }

int shx_alloc(shx_ctx* ctx, size_t bytes, void** dptr)
{
    if (!ctx || !dptr) return SHX_EINVAL;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 8);
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "hipMalloc");
}

int shx_free(shx_ctx* ctx, void* dptr)
{
    if (!ctx) return SHX_EINVAL;
    hipError_t e = hipFree(dptr);
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "hipFree");
}

int shx_h2d(shx_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return SHX_EINVAL;
    if (!bytes) return SHX_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "h2d");
    return sync(ctx);
}

int shx_d2h(shx_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return SHX_EINVAL;
    if (!bytes) return SHX_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "d2h");
    return sync(ctx);
}

int shx_fill(shx_ctx* ctx, double* dst, double value, int64_t n)
{
    if (!ctx) return SHX_EINVAL;
    return launch_fill(ctx, dst, value, n);
}

int shx_copy(shx_ctx* ctx, double* dst, const double* src, int64_t n)
{
    if (!ctx || n < 0 || (n > 0 && (!dst || !src))) return SHX_EINVAL;
    return launch_copy(ctx, dst, src, n);
}

int shx_time_copy(shx_ctx* ctx, double* dst, const double* src, int64_t n, int reps,
                  double* ms_per_copy)
{
    if (!ctx || !ms_per_copy || reps < 1 || n < 1 || !dst || !src) return SHX_EINVAL;
    int rc = launch_copy(ctx, dst, src, n);  // warm-up
    if (rc != SHX_OK) return rc;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return hip_fail(ctx, hipGetLastError(), "time_copy");
    if (hipEventCreate(&b) != hipSuccess) {
        hipEventDestroy(a);
        return hip_fail(ctx, hipGetLastError(), "time_copy");
    }
    hipEventRecord(a, ctx->stream);
    for (int r = 0; r < reps && rc == SHX_OK; ++r) rc = launch_copy(ctx, dst, src, n);
    hipEventRecord(b, ctx->stream);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc != SHX_OK) return rc;
    if (e != hipSuccess) return hip_fail(ctx, e, "time_copy");
    *ms_per_copy = ms / reps;
    return SHX_OK;
}

int shx_derivative_create(shx_ctx* ctx, const shx_derivative_desc* desc, shx_derivative** out)
{
    if (!ctx || !desc || !out) return SHX_EINVAL;
    *out = nullptr;
    const int dir = desc->dir;
    if (dir < 0 || dir > 2) return fail(ctx, SHX_EINVAL, "derivative: dir out of range");
    for (int a = 0; a < 3; ++a)
        if (desc->n[a] < 1 || desc->n[a] > 65535)
            return fail(ctx, SHX_EINVAL, "derivative: extents must be in [1, 65535]");
    if (desc->width != 3 && desc->width != 5)
        return fail(ctx, SHX_EUNSUPPORTED, "derivative: interior width must be 3 or 5");
    const long N = (long)desc->n[0] * desc->n[1] * desc->n[2];
    if (N > 0x7fffffffL) return fail(ctx, SHX_EINVAL, "derivative: mesh too large");

    auto* d = new (std::nothrow) shx_derivative();
    if (!d) return SHX_ENOMEM;
    d->ctx = ctx;
    d->dir = dir;
    for (int a = 0; a < 3; ++a) {
        d->n[a] = desc->n[a];
        d->nR[a] = desc->nR[a];
    }
    std::memset(&d->dev, 0, sizeof d->dev);
    d->dev.dir = dir;
    d->dev.nq = desc->n[dir];
    d->dev.width = desc->width;
    d->dev.stride = dir == 0 ? (long)desc->n[1] * desc->n[2] : dir == 1 ? desc->n[2] : 1;
    const int nslow = dir == 0 ? desc->n[1] : desc->n[0];
    const int nfast = dir == 2 ? desc->n[1] : desc->n[2];
    d->dev.nfast = nfast;
    d->dev.present = desc->n[dir] > 1;
    const long npairs = (long)nslow * nfast;

    // validate the segment table and build pair_ptr and the fast descriptors
    std::vector<int> pair_ptr(npairs + 1, 0);
    std::vector<unsigned long long> fast(npairs, 0ull);
    std::vector<int> seen(npairs, 0);
    long prev_pair = -1;
    int prev_hi = -1;
    for (long b = 0; b < desc->nblocks; ++b) {
        const shx_line_block& blk = desc->blocks[b];
        const int total = blk.nl + blk.ni + blk.nr;
        if (blk.pair < 0 || blk.pair >= npairs || blk.lo < 0 || total < 0 ||
            blk.lo + total > desc->n[dir] || blk.nl < 0 || blk.nr < 0) {
            delete d;
            return fail(ctx, SHX_EINVAL, "derivative: segment out of range");
        }
        const long nRd = desc->nR[dir];
        const bool bad_ref =
            blk.lobj < -1 || blk.lobj >= nRd || blk.robj < -1 || blk.robj >= nRd ||
            blk.lcols < 0 || blk.rcols < 0 || blk.lcoef < 0 || blk.rcoef < 0 ||
            (long)blk.lcoef + (long)blk.nl * blk.lcols > desc->npool ||
            (long)blk.rcoef + (long)blk.nr * blk.rcols > desc->npool ||
            (blk.lobj >= 0 && (blk.lbw < 0 || (long)blk.lbw + blk.nl > desc->npool)) ||
            (blk.robj >= 0 && (blk.rbw < 0 || (long)blk.rbw + blk.nr > desc->npool));
        if (bad_ref) {
            delete d;
            return fail(ctx, SHX_EINVAL, "derivative: segment references out of range");
        }
        if (total < blk.nl || total < blk.nr) {
            delete d;
            return fail(ctx, SHX_EUNSUPPORTED,
                        "derivative: a line segment is shorter than its closures");
        }
        // Overlapping closures of short segments may read past the line ends
        // (into the neighbouring line, as the reference's flat offsets do) but
        // never outside the mesh buffer.
        {
            const int s = blk.pair / nfast, f = blk.pair % nfast;
            long base;
            if (dir == 0) base = (long)s * desc->n[2] + f;
            else if (dir == 1) base = (long)s * desc->n[1] * desc->n[2] + f;
            else base = ((long)s * desc->n[1] + f) * desc->n[2];
            const long st = d->dev.stride;
            const long l0 = base + (long)blk.lcol0 * st, l1 = base + (long)(blk.lcol0 + blk.lcols - 1) * st;
            const long r0 = base + (long)blk.rcol0 * st, r1 = base + (long)(blk.rcol0 + blk.rcols - 1) * st;
            if ((blk.nl > 0 && (l0 < 0 || l1 >= N)) || (blk.nr > 0 && (r0 < 0 || r1 >= N))) {
                delete d;
                return fail(ctx, SHX_EUNSUPPORTED,
                            "derivative: closure columns of a short segment fall outside the mesh");
            }
        }
        if (blk.pair < prev_pair || (blk.pair == prev_pair && blk.lo < prev_hi)) {
            delete d;
            return fail(ctx, SHX_EINVAL, "derivative: segments not sorted by (pair, lo)");
        }
        if (blk.pair != prev_pair) prev_hi = -1;
        prev_pair = blk.pair;
        prev_hi = blk.lo + total;
        pair_ptr[blk.pair + 1]++;
        const int s = seen[blk.pair]++;
        if (s < 2 && blk.ni > 0) {
            const unsigned long long lo = (unsigned long long)(blk.lo + blk.nl);
            const unsigned long long hi = (unsigned long long)(blk.lo + blk.nl + blk.ni);
            fast[blk.pair] |= (lo | (hi << 16)) << (32 * s);
        }
    }
    for (long p = 0; p < npairs; ++p) pair_ptr[p + 1] += pair_ptr[p];

    // every row outside its line's fast intervals, sorted by point: first
    // those that the list evaluates, then the closure rows of the ends that
    // one lane each evaluates (no B term at the end, closures that do not
    // overlap, at most kWallRows rows of kWallCols columns).  The Euler flux
    // derivative takes the whole list.
    std::vector<SlowRow> slow, slow_wall;
    std::vector<WallSide> walls;
    for (long b = 0; b < desc->nblocks; ++b) {
        const shx_line_block& blk = desc->blocks[b];
        const unsigned long long fd = fast[blk.pair];
        const int lo0 = (int)(fd & 0xffffu), hi0 = (int)((fd >> 16) & 0xffffu);
        const int lo1 = (int)((fd >> 32) & 0xffffu), hi1 = (int)(fd >> 48);
        const int s = blk.pair / nfast, f = blk.pair % nfast;
        long base;
        if (dir == 0) base = (long)s * desc->n[2] + f;
        else if (dir == 1) base = (long)s * desc->n[1] * desc->n[2] + f;
        else base = ((long)s * desc->n[1] + f) * desc->n[2];
        const int total = blk.nl + blk.ni + blk.nr;
        const bool lw = blk.nl > 0 && blk.ni >= 0 && blk.lobj < 0 && blk.nl <= kWallRows &&
                        blk.lcols > 0 && blk.lcols <= kWallCols;
        const bool rw = blk.nr > 0 && blk.ni >= 0 && blk.robj < 0 && blk.nr <= kWallRows &&
                        blk.rcols > 0 && blk.rcols <= kWallCols;
        if (lw) walls.push_back(WallSide{base, blk.lcoef, blk.lo, blk.lcol0, (short)blk.nl, (short)blk.lcols});
        if (rw)
            walls.push_back(WallSide{base, blk.rcoef, blk.lo + blk.nl + blk.ni, blk.rcol0, (short)blk.nr,
                                     (short)blk.rcols});
        for (int q = blk.lo; q < blk.lo + total; ++q) {
            if ((q >= lo0 && q < hi0) || (q >= lo1 && q < hi1)) continue;
            const int lr = q - blk.lo;
            const bool end = (lw && lr < blk.nl) || (rw && lr >= blk.nl + blk.ni);
            (end ? slow_wall : slow).push_back(SlowRow{base + (long)q * d->dev.stride, (int)b, q});
        }
    }
    const auto by_idx = [](const SlowRow& a, const SlowRow& b) { return a.idx < b.idx; };
    std::sort(slow.begin(), slow.end(), by_idx);
    std::sort(slow_wall.begin(), slow_wall.end(), by_idx);
    d->nslow_list = (long)slow.size();
    slow.insert(slow.end(), slow_wall.begin(), slow_wall.end());
    d->nslow = (long)slow.size();
    // neighbouring lanes take the same end of neighbouring lines
    std::sort(walls.begin(), walls.end(), [](const WallSide& a, const WallSide& b) {
        return a.q0 != b.q0 ? a.q0 < b.q0 : a.base < b.base;
    });
    d->nwall = (long)walls.size();

    int rc = upload(ctx, &d->pair_ptr, pair_ptr.data(), pair_ptr.size());
    if (rc == SHX_OK && d->nslow) rc = upload(ctx, &d->slow, slow.data(), slow.size());
    if (rc == SHX_OK && d->nwall) rc = upload(ctx, &d->walls, walls.data(), walls.size());
    if (rc == SHX_OK) rc = upload(ctx, &d->fast, fast.data(), fast.size());
    if (rc == SHX_OK && desc->nblocks)
        rc = upload(ctx, &d->blocks, desc->blocks, (size_t)desc->nblocks);
    if (rc == SHX_OK) rc = upload(ctx, &d->pool, desc->pool, (size_t)desc->npool);
    if (rc == SHX_OK) rc = upload(ctx, &d->coef, desc->interior, 9);

    // Neumann entries -> row-compressed CSR sorted by (row, col, w)
    if (rc == SHX_OK && desc->nneu > 0) {
        std::vector<long> order(desc->nneu);
        for (long e = 0; e < desc->nneu; ++e) order[e] = e;
        std::sort(order.begin(), order.end(), [&](long a, long b) {
            if (desc->neu_row[a] != desc->neu_row[b]) return desc->neu_row[a] < desc->neu_row[b];
            if (desc->neu_col[a] != desc->neu_col[b]) return desc->neu_col[a] < desc->neu_col[b];
            return desc->neu_w[a] < desc->neu_w[b];
        });
        std::vector<long> rows, col;
        std::vector<int> ptr{0};
        std::vector<double> w;
        for (long e : order) {
            if (rows.empty() || rows.back() != desc->neu_row[e]) {
                if (!rows.empty()) ptr.push_back((int)w.size());
                rows.push_back(desc->neu_row[e]);
            }
            col.push_back(desc->neu_col[e]);
            w.push_back(desc->neu_w[e]);
        }
        ptr.push_back((int)w.size());
        d->nneu_rows = (long)rows.size();
        rc = upload(ctx, &d->neu_rows, rows.data(), rows.size());
        if (rc == SHX_OK) rc = upload(ctx, &d->neu_ptr, ptr.data(), ptr.size());
        if (rc == SHX_OK) rc = upload(ctx, &d->neu_col, col.data(), col.size());
        if (rc == SHX_OK) rc = upload(ctx, &d->neu_w, w.data(), w.size());
    }
    // R-space rows: (set, row) strictly increasing, columns inside their buffer
    if (rc == SHX_OK && desc->nrrow > 0) {
        const long nr = desc->nrrow;
        bool ok = desc->rset && desc->rrow && desc->rptr && desc->rcol && desc->rw && desc->rptr[0] == 0;
        for (long k = 0; ok && k < nr; ++k) {
            const int set = desc->rset[k];
            ok = set >= 0 && set < 3 && desc->rrow[k] >= 0 && desc->rrow[k] < desc->nR[set] &&
                 desc->rptr[k + 1] >= desc->rptr[k];
            if (ok && k > 0)
                ok = set > desc->rset[k - 1] || (set == desc->rset[k - 1] && desc->rrow[k] > desc->rrow[k - 1]);
            for (long e = desc->rptr[k]; ok && e < desc->rptr[k + 1]; ++e)
                ok = desc->rcol[e] < N && desc->rcol[e] >= -desc->nR[set];
        }
        if (!ok) rc = fail(ctx, SHX_EINVAL, "derivative: R-space rows out of range or unsorted");
        const long nnz = ok ? desc->rptr[nr] : 0;
        d->nrrow = ok ? nr : 0;
        if (rc == SHX_OK) rc = upload(ctx, &d->rset, desc->rset, (size_t)nr);
        if (rc == SHX_OK) rc = upload(ctx, &d->rrow, (const long*)desc->rrow, (size_t)nr);
        if (rc == SHX_OK) rc = upload(ctx, &d->rptr, (const long*)desc->rptr, (size_t)nr + 1);
        if (rc == SHX_OK && nnz) rc = upload(ctx, &d->rcol, (const long*)desc->rcol, (size_t)nnz);
        if (rc == SHX_OK && nnz) rc = upload(ctx, &d->rw, desc->rw, (size_t)nnz);
    }
    if (rc != SHX_OK) {
        shx_derivative_destroy(d);
        return rc;
    }
    d->h_blocks.assign(desc->blocks, desc->blocks + desc->nblocks);
    d->h_pool.assign(desc->pool, desc->pool + desc->npool);
    if (d->nrrow > 0) {
        d->h_rset.assign(desc->rset, desc->rset + d->nrrow);
        d->h_rrow.assign(desc->rrow, desc->rrow + d->nrrow);
        d->h_rptr.assign(desc->rptr, desc->rptr + d->nrrow + 1);
        d->h_rcol.assign(desc->rcol, desc->rcol + d->h_rptr[(size_t)d->nrrow]);
        d->h_rw.assign(desc->rw, desc->rw + d->h_rptr[(size_t)d->nrrow]);
    }
    d->h_pair_ptr = std::move(pair_ptr);
    for (int m = 0; m < 9; ++m) d->h_coef[m] = desc->interior[m];
    d->dev.pair_ptr = d->pair_ptr;
    d->dev.fast = d->fast;
    d->dev.blocks = d->blocks;
    d->dev.pool = d->pool;
    d->dev.coef = d->coef;
    *out = d;
    return SHX_OK;
}

int shx_derivative_destroy(shx_derivative* d)
{
    if (!d) return SHX_EINVAL;
    destroy_transpose(d->T);
    hipFree(d->pair_ptr);
    hipFree(d->fast);
    hipFree(d->slow);
    hipFree(d->walls);
    hipFree(d->blocks);
    hipFree(d->pool);
    hipFree(d->coef);
    hipFree(d->neu_rows);
    hipFree(d->neu_ptr);
    hipFree(d->neu_col);
    hipFree(d->neu_w);
    hipFree(d->rset);
    hipFree(d->rrow);
    hipFree(d->rptr);
    hipFree(d->rcol);
    hipFree(d->rw);
    delete d;
    return SHX_OK;
}

int shx_derivative_apply(shx_derivative* d, const shx_field* u, const shx_field* nu, shx_field* du,
                         int accumulate)
{
    if (!d || !u || !du) return SHX_EINVAL;
    int rc = launch_cut_rows(d, u, du);
    if (rc == SHX_OK) rc = launch_derivative(d, u->D, u->R[d->dir], du->D, accumulate != 0, true);
    if (rc == SHX_OK && nu) rc = launch_neumann(d, nu->D, du->D);
    if (rc == SHX_OK) rc = sync(d->ctx);
    return rc;
}

int shx_time_derivative_apply(shx_derivative* d, const shx_field* u, shx_field* du, int accumulate,
                              int reps, double* ms_per_apply)
{
    if (!d || !u || !du || reps < 1 || !ms_per_apply) return SHX_EINVAL;
    int rc = launch_derivative(d, u->D, u->R[d->dir], du->D, accumulate != 0, true);  // warm-up
    if (rc == SHX_OK) rc = sync(d->ctx);
    if (rc != SHX_OK) return rc;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return hip_fail(d->ctx, hipGetLastError(), "event");
    if (hipEventCreate(&b) != hipSuccess) {
        hipEventDestroy(a);
        return hip_fail(d->ctx, hipGetLastError(), "event");
    }
    hipEventRecord(a, d->ctx->stream);
    for (int r = 0; r < reps && rc == SHX_OK; ++r)
        rc = launch_derivative(d, u->D, u->R[d->dir], du->D, accumulate != 0, true);
    hipEventRecord(b, d->ctx->stream);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc != SHX_OK) return rc;
    if (e != hipSuccess) return hip_fail(d->ctx, e, "time derivative");
    *ms_per_apply = ms / reps;
    return SHX_OK;
}

int shx_block_apply(shx_derivative* d, const double* x, double* b, int accumulate)
{
    if (!d || !x || !b) return SHX_EINVAL;
    int rc = launch_derivative(d, x, nullptr, b, accumulate != 0, false);
    if (rc == SHX_OK) rc = sync(d->ctx);
    return rc;
}

int shx_csr_create(shx_ctx* ctx, int64_t rows, const int64_t* rowptr, const int64_t* col,
                   const double* w, shx_csr** out)
{
    if (!ctx || !out || rows < 0 || (rows > 0 && !rowptr)) return SHX_EINVAL;
    auto* A = new (std::nothrow) shx_csr();
    if (!A) return SHX_ENOMEM;
    A->ctx = ctx;
    A->rows = rows;
    const long nnz = rows > 0 ? (long)rowptr[rows] : 0;
    A->nnz = nnz;
    int rc = upload(ctx, &A->u, (const long*)rowptr, rows > 0 ? (size_t)rows + 1 : 0);
    if (rc == SHX_OK) rc = upload(ctx, &A->v, (const long*)col, (size_t)nnz);
    if (rc == SHX_OK) rc = upload(ctx, &A->w, w, (size_t)nnz);
    if (rc != SHX_OK) {
        shx_csr_destroy(A);
        return rc;
    }
    *out = A;
    return SHX_OK;
}

int shx_csr_destroy(shx_csr* A)
{
    if (!A) return SHX_EINVAL;
    hipFree(A->u);
    hipFree(A->v);
    hipFree(A->w);
    delete A;
    return SHX_OK;
}

int shx_csr_apply(shx_csr* A, const double* x, double* b)
{
    if (!A) return SHX_EINVAL;
    int rc = launch_csr_full(A, x, b);
    if (rc == SHX_OK) rc = sync(A->ctx);
    return rc;
}

int shx_gradient_apply(shx_derivative* const d[3], const shx_field* u, shx_field* const du[3])
{
    shx_derivative* any = d[0] ? d[0] : d[1] ? d[1] : d[2];
    if (!any || !u || !du[0] || !du[1] || !du[2]) return SHX_EINVAL;
    shx_ctx* ctx = any->ctx;
    int rc = SHX_OK;
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) rc = zero_field(ctx, du[c], any->n, any->nR);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c)
        if (d[c] && d[c]->n[c] > 1) {
            rc = launch_cut_rows(d[c], u, du[c]);
            if (rc == SHX_OK) rc = launch_derivative(d[c], u->D, u->R[c], du[c]->D, false, true);
        }
    if (rc == SHX_OK) rc = sync(ctx);
    return rc;
}

int shx_laplacian_apply(shx_derivative* const d[3], const shx_field* u, const shx_field* nu,
                        shx_field* du)
{
    shx_derivative* any = d[0] ? d[0] : d[1] ? d[1] : d[2];
    if (!any || !u || !du) return SHX_EINVAL;
    shx_ctx* ctx = any->ctx;
    int rc = zero_field(ctx, du, any->n, any->nR);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
        if (!d[c] || d[c]->n[c] < 2) continue;
        rc = launch_cut_rows(d[c], u, du);
        if (rc == SHX_OK) rc = launch_derivative(d[c], u->D, u->R[c], du->D, true, true);
        if (rc == SHX_OK && nu) rc = launch_neumann(d[c], nu->D, du->D);
    }
    if (rc == SHX_OK) rc = sync(ctx);
    return rc;
}

}  // extern "C"
