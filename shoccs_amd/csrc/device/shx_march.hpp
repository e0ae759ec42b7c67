// shx_march.hpp -- device templates of the x-marching form of one fused RK4
// stage, instantiated by shx_march.hip (FP64 form) and shx_mixed.hip (mixed
// form: FP32 stage register).
//
// A workgroup of the tile part owns a (y, z) tile of TY x 64 points and walks
// a chunk of x planes.  Each thread keeps the x-neighbours of its point in
// registers (a queue shifted by one plane per step); the current plane is
// staged in LDS, double-buffered, with W/2-point y and z halos, so each stage
// value is fetched about once and there is one barrier per plane.  Interior,
// solid and closure-pattern points are evaluated there; the remaining points
// (rows with B terms, closures out of the tile's reach, Dirichlet faces) come
// from records, evaluated by the special part, which also sets the Dirichlet
// cut points.  The FP64 form uses the formulas of shx_point.hpp in the
// reference's order, so its results are bit-identical to the point-wise
// tiles.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "shx_point.hpp"
#include "shx_sys.hpp"

namespace shx {

namespace {

// Interior circulant row on FP32 values (the mixed form's stage register):
// summed in FP32 with Kahan compensation.  (FP64 values are summed in FP64
// from 0.0, left to right, the reference's order, inline in march_tile.)
template <int W>
__device__ __forceinline__ double idot(const float (&c)[W], const float (&v)[W])
{
    float sum = 0.f, comp = 0.f;
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const float y = c[m] * v[m] - comp;
        const float t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    return (double)sum;
}

// A y or z closure row from the LDS tile, in FP64: `c0` is the tile slot of
// the point, `step` the slot distance of one position along the direction.
template <typename TI>
__device__ __forceinline__ double tile_row(const ClosureRow& r, const TI* c0, int step)
{
    double s = 0.0;
#pragma unroll 1
    for (int m = 0; m < r.n; ++m) s += r.w[m] * (double)c0[(r.off + m) * step];
    return s;
}

// Stage register types: the FP64 form keeps every register in FP64; the
// mixed form keeps the stage register in FP32 (stage 0 reads u0, stage 3
// writes u0, both FP64).
template <bool MIXED, int STAGE>
struct StageTypes {
    using in = typename std::conditional<MIXED && STAGE != 0, float, double>::type;
    using out = typename std::conditional<MIXED && STAGE != 3, float, double>::type;
};

// One march tile: logical workgroup L of the tile grid.
template <int KIND, int STAGE, int W, int TY, bool MIXED>
__device__ __forceinline__ void march_tile(const MarchArgs& a, const StageScalars& ss, unsigned L)
{
    using TI = typename StageTypes<MIXED, STAGE>::in;
    using TO = typename StageTypes<MIXED, STAGE>::out;
    constexpr int TZ = 64;
    constexpr int hw = W / 2;
    constexpr int LY = TY + 2 * hw, LZ = TZ + 2 * hw;
    constexpr int HQ = hw + kXReach;  // x queue half width
    constexpr int QW = 2 * HQ + 1;    // x queue length
    static_assert(TY >= 2 * hw, "y halo rows are loaded by distinct threads");
    __shared__ TI tile[2][LY][LZ];
    __shared__ ClosureRow crow[3][kMaxClosureRows];

    // tile -> (z tile, y tile, x chunk), z fastest
    const int bz = (int)(L % (unsigned)a.nbz);
    const unsigned r1 = L / (unsigned)a.nbz;
    const int by = (int)(r1 % (unsigned)a.nby);
    const int bx = (int)(r1 / (unsigned)a.nby);
    const int tz = threadIdx.x % TZ, ty = threadIdx.x / TZ;
    const int n0 = a.s.n0, n1 = a.s.n1, n2 = a.s.n2;
    const int k = bz * TZ + tz, j = by * TY + ty;
    const bool active = k < n2 && j < n1;
    const int kc = min(k, n2 - 1), jc = min(j, n1 - 1);
    const long plane = (long)n1 * n2;
    const long col = (long)jc * n2 + kc;

    // halo sources inside a plane (positions clamped to the mesh; clamped
    // values are never used by a row evaluated here)
    int yslot = -1, zslot = -1;
    long yoff = 0, zoff = 0;
    if (ty < hw) {
        const int jj = min(max(by * TY - hw + ty, 0), n1 - 1);
        yslot = ty;
        yoff = (long)jj * n2 + kc;
    } else if (ty >= TY - hw) {
        const int jj = min(by * TY + TY + ty - (TY - hw), n1 - 1);
        yslot = ty + 2 * hw;
        yoff = (long)jj * n2 + kc;
    }
    if (tz < hw) {
        const int kk = min(max(bz * TZ - hw + tz, 0), n2 - 1);
        zslot = tz;
        zoff = (long)jc * n2 + kk;
    } else if (tz >= TZ - hw) {
        const int kk = min(bz * TZ + TZ + tz - (TZ - hw), n2 - 1);
        zslot = tz + 2 * hw;
        zoff = (long)jc * n2 + kk;
    }

    // closure-row tables to LDS (read after the first barrier of the march)
    {
        const int nwords = 3 * kMaxClosureRows * (int)(sizeof(ClosureRow) / sizeof(double));
        const double* src = reinterpret_cast<const double*>(a.rows);
        double* dst = reinterpret_cast<double*>(&crow[0][0]);
        for (int w = threadIdx.x; w < nwords; w += 64 * TY) dst[w] = src[w];
    }
    const int ny1 = a.nrows[1] + 1, nz1 = a.nrows[2] + 1;

    const bool px = a.ops.d[0].present, py = a.ops.d[1].present, pz = a.ops.d[2].present;
    TI cx[W], cy[W], cz[W];
#pragma unroll
    for (int m = 0; m < W; ++m) {
        cx[m] = (TI)(px ? a.ops.d[0].coef[m] : 0.0);
        cy[m] = (TI)(py ? a.ops.d[1].coef[m] : 0.0);
        cz[m] = (TI)(pz ? a.ops.d[2].coef[m] : 0.0);
    }

    const int x0 = bx * a.xchunk;
    const int x1 = min(x0 + a.xchunk, n0);
    const TI* __restrict__ us = static_cast<const TI*>(a.us);
    TO* out = static_cast<TO*>(a.out);
    // x queue: q[m] = us at plane x - HQ + m; qn is the next queue element.
    TI q[QW];
#pragma unroll
    for (int m = 0; m < QW - 1; ++m) q[m] = us[(long)min(max(x0 - HQ + m, 0), n0 - 1) * plane + col];
    TI qn = us[(long)min(x0 + HQ, n0 - 1) * plane + col];
    // values of the current plane; those of the next plane are loaded one
    // iteration ahead so that their latency overlaps the current plane
    long pcur = (long)x0 * plane;
    TI yv = yslot >= 0 ? us[pcur + yoff] : (TI)0;
    TI zv = zslot >= 0 ? us[pcur + zoff] : (TI)0;
    unsigned char c = a.cls[pcur + col];
    double u0v = a.u0[pcur + col];
    double accv = STAGE == 0 ? 0.0 : a.acc[pcur + col];

    for (int x = x0; x < x1; ++x) {
        const long pb = (long)x * plane;
        q[QW - 1] = qn;
        // next plane (the last iteration reloads its own plane; unused)
        const int xn = min(x + 1, x1 - 1);
        const long pn = (long)xn * plane;
        qn = us[(long)min(xn + HQ, n0 - 1) * plane + col];
        const TI nyv = yslot >= 0 ? us[pn + yoff] : (TI)0;
        const TI nzv = zslot >= 0 ? us[pn + zoff] : (TI)0;
        // single-use streams: non-temporal, so that L2 keeps the stage
        // values that neighbouring tiles and chunks re-read
        const unsigned char nc = __builtin_nontemporal_load(a.cls + pn + col);
        const double nu0 = __builtin_nontemporal_load(a.u0 + pn + col);
        const double nacc = STAGE == 0 ? 0.0 : __builtin_nontemporal_load(a.acc + pn + col);
        const long idx = pb + col;
        TI(*t)[LZ] = tile[x & 1];
        t[ty + hw][tz + hw] = q[HQ];
        if (yslot >= 0) t[yslot][tz + hw] = yv;
        if (zslot >= 0) t[ty + hw][zslot] = zv;
        __syncthreads();
        if (active && c != kSpecial) {
            double rhs = 0.0;
            double v;
            {
                if (c != kNoRows) {
                    int ix = 0, iy = 0, iz = 0;
                    if (c >= kPattern0) {
                        int p = c - kPattern0;
                        iz = p % nz1;
                        p /= nz1;
                        iy = p % ny1;
                        ix = p / ny1;
                    }
                    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
                    if (px) {
                        double sx = 0.0;
                        if (ix == 0) {
                            if constexpr (std::is_same<TI, double>::value) {
#pragma unroll
                                for (int m = 0; m < W; ++m) sx += cx[m] * q[kXReach + m];
                            } else {
                                TI v[W];
#pragma unroll
                                for (int m = 0; m < W; ++m) v[m] = q[kXReach + m];
                                sx = idot<W>(cx, v);
                            }
                        } else {
                            // closure row over the queue: only its own columns,
                            // in column order
                            const ClosureRow& r = crow[0][ix - 1];
#pragma unroll
                            for (int m = 0; m < QW; ++m) {
                                const int ci = m - HQ - r.off;
                                if (ci >= 0 && ci < r.n) sx += r.w[ci] * (double)q[m];
                            }
                        }
                        d0 = sx;
                    }
                    if (py) {
                        if (iy == 0) {
                            if constexpr (std::is_same<TI, double>::value) {
                                double sy = 0.0;
#pragma unroll
                                for (int m = 0; m < W; ++m) sy += cy[m] * t[ty + m][tz + hw];
                                d1 = sy;
                            } else {
                                TI v[W];
#pragma unroll
                                for (int m = 0; m < W; ++m) v[m] = t[ty + m][tz + hw];
                                d1 = idot<W>(cy, v);
                            }
                        } else {
                            d1 = tile_row(crow[1][iy - 1], &t[ty + hw][tz + hw], LZ);
                        }
                    }
                    if (pz) {
                        if (iz == 0) {
                            if constexpr (std::is_same<TI, double>::value) {
                                double sz = 0.0;
#pragma unroll
                                for (int m = 0; m < W; ++m) sz += cz[m] * t[ty + hw][tz + m];
                                d2 = sz;
                            } else {
                                TI v[W];
#pragma unroll
                                for (int m = 0; m < W; ++m) v[m] = t[ty + hw][tz + m];
                                d2 = idot<W>(cz, v);
                            }
                        } else {
                            d2 = tile_row(crow[2][iz - 1], &t[ty + hw][tz + hw], 1);
                        }
                    }
                    if (kIsHeat<KIND>) {
                        // laplacian rows added in direction order (plus_eq)
                        double lap = 0.0;
                        if (px) lap += d0;
                        if (py) lap += d1;
                        if (pz) lap += d2;
                        rhs = heat_finish<KIND>(a.s, false, true, x, j, k, lap, ss.t_s);
                    } else {
                        rhs = wave_combine<KIND>(a.s, false, x, j, k, d0, d1, d2);
                    }
                } else {
                    rhs = kIsHeat<KIND> ? heat_finish<KIND>(a.s, false, false, x, j, k, 0.0, ss.t_s)
                                               : wave_combine<KIND>(a.s, false, x, j, k, 0.0, 0.0, 0.0);
                }
                // stage_update, with non-temporal stores
                const double aa = (STAGE == 0 ? 0.0 : accv) + ss.c_acc * rhs;
                if (STAGE < 3) {
                    __builtin_nontemporal_store(aa, a.acc + idx);
                    v = u0v + ss.c_next * rhs;
                } else {
                    v = u0v + 1.0 * aa;
                }
            }
            __builtin_nontemporal_store((TO)v, out + idx);
        }
#pragma unroll
        for (int m = 0; m < QW - 1; ++m) q[m] = q[m + 1];
        yv = nyv;
        zv = nzv;
        c = nc;
        u0v = nu0;
        accv = nacc;
    }
}

// Row of direction d at a special point from its record (the same sums as
// row_at: interior rows tap by tap, other rows as listed, B terms after).
template <int W, typename TI>
__device__ __forceinline__ RowVal special_row(const DevDeriv& d, const SpecialRow* __restrict__ rows,
                                              unsigned ref, const TI* __restrict__ u,
                                              const double* __restrict__ uR, long idx)
{
    RowVal r;
    r.dot = 0.0;
    r.b0 = r.b1 = 0.0;
    r.nb = 0;
    r.row = ref != kRowNone;
    if (!r.row) return r;
    const long st = d.stride;
    double dot = 0.0;
    if (ref == kRowInterior) {
        constexpr int hw = W / 2;
#pragma unroll
        for (int m = 0; m < W; ++m) dot += d.coef[m] * (double)u[idx + (long)(m - hw) * st];
        r.dot = dot;
        return r;
    }
    const SpecialRow e = rows[ref];
    const double* __restrict__ w = d.pool + e.coef;
    const TI* __restrict__ v = u + idx + (long)e.off * st;
    for (int c = 0; c < e.ncols; ++c) dot += w[c] * (double)v[(long)c * st];
    r.dot = dot;
    r.nb = e.nb;
    if (e.nb > 0) r.b0 = d.pool[e.bw0] * uR[e.bobj0];
    if (e.nb > 1) r.b1 = d.pool[e.bw1] * uR[e.bobj1];
    return r;
}

// Item q of the special work: a listed point, then the cut points.  Rows are
// evaluated in FP64 in both forms.
template <int KIND, int STAGE, int W, bool MIXED>
__device__ __forceinline__ void special_item(const SpecialArgs& a, const StageScalars& ss, long q)
{
    using TI = typename StageTypes<MIXED, STAGE>::in;
    using TO = typename StageTypes<MIXED, STAGE>::out;
    const TI* us = static_cast<const TI*>(a.us);
    if (q < a.nspecial) {
        const SpecialRec rec = a.recs[q];
        const long idx = rec.idx;
        int i, j, k;
        point_coords(idx, a.s.n1, a.s.n2, i, j, k);
        const DevOps& ops = a.ops;
        double rhs;
        if (kIsHeat<KIND>) {
            double lap = 0.0;
            if (ops.d[0].present) {
                const RowVal r = special_row<W>(ops.d[0], a.rows0, rec.ref[0], us, a.usR.r[0], idx);
                if (r.row) lap = row_plus(lap, r);
            }
            if (ops.d[1].present) {
                const RowVal r = special_row<W>(ops.d[1], a.rows1, rec.ref[1], us, a.usR.r[1], idx);
                if (r.row) lap = row_plus(lap, r);
            }
            if (ops.d[2].present) {
                const RowVal r = special_row<W>(ops.d[2], a.rows2, rec.ref[2], us, a.usR.r[2], idx);
                if (r.row) lap = row_plus(lap, r);
            }
            rhs = heat_finish<KIND>(a.s, on_dirichlet_face(a.s, i, j, k), is_fluid(a.s, idx), i, j, k, lap,
                              ss.t_s);
        } else {
            double du0 = 0.0, du1 = 0.0, du2 = 0.0;
            if (ops.d[0].present) {
                const RowVal r = special_row<W>(ops.d[0], a.rows0, rec.ref[0], us, a.usR.r[0], idx);
                if (r.row) du0 = row_eq(r);
            }
            if (ops.d[1].present) {
                const RowVal r = special_row<W>(ops.d[1], a.rows1, rec.ref[1], us, a.usR.r[1], idx);
                if (r.row) du1 = row_eq(r);
            }
            if (ops.d[2].present) {
                const RowVal r = special_row<W>(ops.d[2], a.rows2, rec.ref[2], us, a.usR.r[2], idx);
                if (r.row) du2 = row_eq(r);
            }
            rhs = wave_combine<KIND>(a.s, on_dirichlet_face(a.s, i, j, k), i, j, k, du0, du1, du2);
        }
        stage_store<KIND, STAGE>(a.s, rhs, a.u0, a.acc, static_cast<TO*>(a.out), idx, i, j, k, ss.c_acc,
                                 ss.c_next, ss.t_next);
        return;
    }
    q -= a.nspecial;
    if (q < a.nR0) {
        if (a.dir0[q])
            a.outR0[q] = exact_of<KIND>(a.s, a.pos0[3 * q], a.pos0[3 * q + 1], a.pos0[3 * q + 2], ss.t_next);
        return;
    }
    q -= a.nR0;
    if (q < a.nR1) {
        if (a.dir1[q])
            a.outR1[q] = exact_of<KIND>(a.s, a.pos1[3 * q], a.pos1[3 * q + 1], a.pos1[3 * q + 2], ss.t_next);
        return;
    }
    q -= a.nR1;
    if (q < a.nR2) {
        if (a.dir2[q])
            a.outR2[q] = exact_of<KIND>(a.s, a.pos2[3 * q], a.pos2[3 * q + 1], a.pos2[3 * q + 2], ss.t_next);
    }
}

// One fused RK4 stage.  Workgroups go in groups of 8 (one per XCD): `sg`
// groups evaluate the special items and `mg` groups are march tiles.  The
// two parts read only the stage input and write disjoint points, so any
// placement is correct; `mode` picks it: 1 special groups first, 2 last,
// 3 spread evenly (group g is special when floor((g+1) sg / T) exceeds
// floor(g sg / T), T = sg + mg).  A launch with sg = 0 or mg = 0 runs one
// part alone.  Tile workgroups keep their XCD (b mod 8) and map to tiles
// contiguously per XCD.
template <int KIND, int STAGE, int W, int TY, bool MIXED>
__device__ __forceinline__ void stage_block(const MarchArgs& a, const SpecialArgs& sp,
                                            const StageScalars& ss, unsigned b, unsigned sg,
                                            unsigned mg, int mode)
{
    const unsigned g = b / 8u, lane8 = b % 8u;
    bool special;
    unsigned si, mi;  // group index within its part
    if (mode == 1) {
        special = g < sg;
        si = g;
        mi = g - sg;
    } else if (mode == 2) {
        special = g >= mg;
        si = g - mg;
        mi = g;
    } else {
        const unsigned long long T = (unsigned long long)sg + mg;
        const unsigned before = (unsigned)((unsigned long long)g * sg / T);
        const unsigned upto = (unsigned)((unsigned long long)(g + 1) * sg / T);
        special = upto > before;
        si = before;
        mi = g - before;
    }
    if (special) {
        const long blk = (long)si * 8 + lane8;
        special_item<KIND, STAGE, W, MIXED>(sp, ss, blk * (64 * TY) + threadIdx.x);
        return;
    }
    const unsigned m = mi * 8u + lane8;
    if (m >= a.nblocks) return;
    march_tile<KIND, STAGE, W, TY, MIXED>(a, ss, xcd_remap(m, a.nblocks));
}

template <int KIND, int STAGE, int W, int TY, bool MIXED>
__global__ __launch_bounds__(64 * TY) void k_stage(MarchArgs a, SpecialArgs sp, StageScalars ss,
                                                   unsigned sg, unsigned mg, int mode)
{
    stage_block<KIND, STAGE, W, TY, MIXED>(a, sp, ss, blockIdx.x, sg, mg, mode);
}

// RK4 stage constants (rk4.cpp:13-14), as the host's
constexpr double kStageTime[4] = {0.0, 0.5, 0.5, 1.0};
constexpr double kStageWeight[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

// One fused RK4 stage of a batch of problems in one launch: problem p owns
// workgroups [start[p], start[p+1]) (multiples of 8, so that tile workgroups
// keep their XCD); its arguments and clock live in device memory, so that a
// captured launch stays valid at any time.  The stage scalars are formed as
// the single-problem host path forms them (t = t0 + n dt).
template <int KIND, int STAGE, int W, int TY, bool MIXED>
__global__ __launch_bounds__(64 * TY) void k_batch_stage(const BatchEntry* __restrict__ ent,
                                                         const BatchClock* __restrict__ clk,
                                                         const unsigned* __restrict__ start,
                                                         int nprob)
{
    int lo = 0, hi = nprob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (start[mid] <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    // the problem's arguments: read-only for the kernel's lifetime, read
    // through the constant address space so that their uniform loads go to
    // scalar registers and may be repeated instead of kept live
#if defined(__HIP_DEVICE_COMPILE__)
    using CEntry = const __attribute__((address_space(4))) BatchEntry;
    const CEntry* ce = (const CEntry*)ent;
#else  // the host pass only parses the kernel
    const BatchEntry* ce = ent;
#endif
    const MarchArgs m = ce[lo].m;
    const SpecialArgs sp = ce[lo].sp;
    const unsigned sg = ce[lo].sg, mg = ce[lo].mg;
    const int mode = ce[lo].mode;
    const double dt = clk[lo].dt;
    const double t = clk[lo].t0 + (double)clk[lo].n * dt;
    StageScalars ss;
    ss.t_s = t + dt * kStageTime[STAGE];
    ss.c_acc = dt * kStageWeight[STAGE];
    ss.c_next = STAGE < 3 ? dt * kStageTime[STAGE + 1] : 1.0;
    ss.t_next = STAGE < 3 ? t + dt * kStageTime[STAGE + 1] : t + dt;
    stage_block<KIND, STAGE, W, TY, MIXED>(m, sp, ss, blockIdx.x - start[lo], sg, mg, mode);
}

template <int KIND, int STAGE, int W, bool MIXED>
int stage_ty(hipStream_t st, int ty, int mode, const MarchArgs& a, const SpecialArgs& sp,
             const StageScalars& ss, long nitems)
{
    const int threads = 64 * ty;
    const unsigned sg = (unsigned)(((nitems + threads - 1) / threads + 7) / 8);
    const unsigned mg = (a.nblocks + 7u) / 8u;
    auto go = [&](unsigned s_groups, unsigned m_groups, int md) {
        const unsigned grid = 8u * (s_groups + m_groups);
        if (grid == 0) return 0;
        switch (ty) {
        case 4: k_stage<KIND, STAGE, W, 4, MIXED><<<grid, 64 * 4, 0, st>>>(a, sp, ss, s_groups, m_groups, md); return 0;
        case 8: k_stage<KIND, STAGE, W, 8, MIXED><<<grid, 64 * 8, 0, st>>>(a, sp, ss, s_groups, m_groups, md); return 0;
        case 16: k_stage<KIND, STAGE, W, 16, MIXED><<<grid, 64 * 16, 0, st>>>(a, sp, ss, s_groups, m_groups, md); return 0;
        default: return -1;
        }
    };
    if (mode == 0) {  // two launches: tiles, then special items
        const int rc = go(0, mg, 1);
        return rc != 0 ? rc : go(sg, 0, 1);
    }
    return go(sg, mg, mode);
}

template <int KIND, int W, bool MIXED>
int stage_of(hipStream_t st, int stage, int ty, int mode, const MarchArgs& a, const SpecialArgs& sp,
             const StageScalars& ss, long nitems)
{
    switch (stage) {
    case 0: return stage_ty<KIND, 0, W, MIXED>(st, ty, mode, a, sp, ss, nitems);
    case 1: return stage_ty<KIND, 1, W, MIXED>(st, ty, mode, a, sp, ss, nitems);
    case 2: return stage_ty<KIND, 2, W, MIXED>(st, ty, mode, a, sp, ss, nitems);
    case 3: return stage_ty<KIND, 3, W, MIXED>(st, ty, mode, a, sp, ss, nitems);
    default: return -1;
    }
}

// One fused RK4 stage of the x-marching form (FP64 or mixed).
template <bool MIXED>
int launch_stage_form(hipStream_t st, int kind, int stage, int width, int ty, int mode,
                      const MarchArgs& a, const SpecialArgs& sp, const StageScalars& ss)
{
    const long nitems = sp.nspecial + sp.nR0 + sp.nR1 + sp.nR2;
    if (a.nblocks == 0 && nitems == 0) return 0;
    switch (kind * 10 + width) {
    case SHX_SYS_HEAT * 10 + 5: return stage_of<SHX_SYS_HEAT, 5, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    case SHX_SYS_HEAT * 10 + 3: return stage_of<SHX_SYS_HEAT, 3, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    case kHeatGauss * 10 + 5: return stage_of<kHeatGauss, 5, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    case kHeatGauss * 10 + 3: return stage_of<kHeatGauss, 3, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    case SHX_SYS_WAVE_LINEAR * 10 + 5:
        return stage_of<SHX_SYS_WAVE_LINEAR, 5, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    case SHX_SYS_WAVE_LINEAR * 10 + 3:
        return stage_of<SHX_SYS_WAVE_LINEAR, 3, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    case SHX_SYS_WAVE_RADIAL * 10 + 5:
        return stage_of<SHX_SYS_WAVE_RADIAL, 5, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    case SHX_SYS_WAVE_RADIAL * 10 + 3:
        return stage_of<SHX_SYS_WAVE_RADIAL, 3, MIXED>(st, stage, ty, mode, a, sp, ss, nitems);
    default: return -1;
    }
}


template <int KIND, int STAGE, int W, bool MIXED>
int batch_ty(hipStream_t st, int ty, const BatchLaunch& bl)
{
    switch (ty) {
    case 4:
        k_batch_stage<KIND, STAGE, W, 4, MIXED><<<bl.grid, 64 * 4, 0, st>>>(bl.ent, bl.clk, bl.start, bl.nprob);
        return 0;
    case 8:
        k_batch_stage<KIND, STAGE, W, 8, MIXED><<<bl.grid, 64 * 8, 0, st>>>(bl.ent, bl.clk, bl.start, bl.nprob);
        return 0;
    case 16:
        k_batch_stage<KIND, STAGE, W, 16, MIXED><<<bl.grid, 64 * 16, 0, st>>>(bl.ent, bl.clk, bl.start, bl.nprob);
        return 0;
    default:
        return -1;
    }
}

template <int KIND, int W, bool MIXED>
int batch_stage_of(hipStream_t st, int stage, int ty, const BatchLaunch& bl)
{
    switch (stage) {
    case 0: return batch_ty<KIND, 0, W, MIXED>(st, ty, bl);
    case 1: return batch_ty<KIND, 1, W, MIXED>(st, ty, bl);
    case 2: return batch_ty<KIND, 2, W, MIXED>(st, ty, bl);
    case 3: return batch_ty<KIND, 3, W, MIXED>(st, ty, bl);
    default: return -1;
    }
}

// One fused RK4 stage of a batch (FP64 or mixed).
template <bool MIXED>
int launch_batch_form(hipStream_t st, int kind, int stage, int width, int ty, const BatchLaunch& bl)
{
    if (bl.grid == 0) return 0;
    switch (kind * 10 + width) {
    case SHX_SYS_HEAT * 10 + 5: return batch_stage_of<SHX_SYS_HEAT, 5, MIXED>(st, stage, ty, bl);
    case SHX_SYS_HEAT * 10 + 3: return batch_stage_of<SHX_SYS_HEAT, 3, MIXED>(st, stage, ty, bl);
    case kHeatGauss * 10 + 5: return batch_stage_of<kHeatGauss, 5, MIXED>(st, stage, ty, bl);
    case kHeatGauss * 10 + 3: return batch_stage_of<kHeatGauss, 3, MIXED>(st, stage, ty, bl);
    case SHX_SYS_WAVE_LINEAR * 10 + 5: return batch_stage_of<SHX_SYS_WAVE_LINEAR, 5, MIXED>(st, stage, ty, bl);
    case SHX_SYS_WAVE_LINEAR * 10 + 3: return batch_stage_of<SHX_SYS_WAVE_LINEAR, 3, MIXED>(st, stage, ty, bl);
    case SHX_SYS_WAVE_RADIAL * 10 + 5: return batch_stage_of<SHX_SYS_WAVE_RADIAL, 5, MIXED>(st, stage, ty, bl);
    case SHX_SYS_WAVE_RADIAL * 10 + 3: return batch_stage_of<SHX_SYS_WAVE_RADIAL, 3, MIXED>(st, stage, ty, bl);
    default: return -1;
    }
}

}  // namespace

}  // namespace shx
