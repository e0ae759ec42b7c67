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

// Planes of loads a march tile keeps in flight ahead of the plane it
// evaluates (a ring of per-plane load slots in registers).
#ifndef SHX_MARCH_PF
#define SHX_MARCH_PF 2
#endif
constexpr int kMarchAhead = SHX_MARCH_PF;

// Vector of V elements of T (one 8- or 16-byte access for V = 2).
template <typename T, int V>
struct VecOf {
    typedef T type __attribute__((ext_vector_type(V)));
};

// VEC consecutive elements at p (p aligned to VEC elements): plain and
// non-temporal loads, non-temporal store.
template <int VEC, typename T>
__device__ __forceinline__ void vload(const T* p, T (&o)[VEC])
{
    if constexpr (VEC == 1) {
        o[0] = *p;
    } else {
        const auto v = *reinterpret_cast<const typename VecOf<T, VEC>::type*>(p);
#pragma unroll
        for (int e = 0; e < VEC; ++e) o[e] = v[e];
    }
}
template <int VEC, typename T>
__device__ __forceinline__ void vload_nt(const T* p, T (&o)[VEC])
{
    if constexpr (VEC == 1) {
        o[0] = __builtin_nontemporal_load(p);
    } else {
        const auto v = __builtin_nontemporal_load(reinterpret_cast<const typename VecOf<T, VEC>::type*>(p));
#pragma unroll
        for (int e = 0; e < VEC; ++e) o[e] = v[e];
    }
}
template <int VEC, typename T>
__device__ __forceinline__ void vstore_nt(T* p, const T (&x)[VEC])
{
    if constexpr (VEC == 1) {
        __builtin_nontemporal_store(x[0], p);
    } else {
        typename VecOf<T, VEC>::type v;
#pragma unroll
        for (int e = 0; e < VEC; ++e) v[e] = x[e];
        __builtin_nontemporal_store(v, reinterpret_cast<typename VecOf<T, VEC>::type*>(p));
    }
}
// VEC class bytes (VEC = 1 or 2) at p, non-temporal.
template <int VEC>
__device__ __forceinline__ void vload_cls(const unsigned char* p, unsigned char (&o)[VEC])
{
    static_assert(VEC == 1 || VEC == 2, "class bytes come one or two per lane");
    if constexpr (VEC == 1) {
        o[0] = __builtin_nontemporal_load(p);
    } else {
        const unsigned short w = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p));
        o[0] = (unsigned char)(w & 0xffu);
        o[1] = (unsigned char)(w >> 8);
    }
}

// What a lane loads for one plane of its march: the x-queue entries (stage
// values HQ planes ahead), its y halo values (VEC) and z halo value, and the
// class bytes, u0 and accumulator of its points.
template <typename TI, int VEC>
struct PlaneLoad {
    TI q[VEC], y[VEC];
    TI z;
    double u0[VEC], acc[VEC];
    unsigned char c[VEC];
};

// One march tile: logical workgroup L of the tile grid.  A lane owns VEC
// consecutive points along z (VEC = 2 needs an even z extent, so that the
// pairs are aligned and never straddle a row end); a tile row holds 64 VEC
// points.
template <int KIND, int STAGE, int W, int TY, int VEC, bool MIXED>
__device__ __forceinline__ void march_tile(const MarchArgs& a, const StageScalars& ss, unsigned L)
{
    using TI = typename StageTypes<MIXED, STAGE>::in;
    using TO = typename StageTypes<MIXED, STAGE>::out;
    constexpr int TZ = 64 * VEC;
    constexpr int hw = W / 2;
    constexpr int LY = TY + 2 * hw, LZ = TZ + 2 * hw;
    constexpr int HQ = hw + kXReach;  // x queue half width
    constexpr int QW = 2 * HQ + 1;    // x queue length
    static_assert(TY >= 2 * hw, "y halo rows are loaded by distinct threads");
    __shared__ TI tile[2][LY][LZ];
    __shared__ ClosureRow crow[3][kMaxClosureRows];
    __shared__ unsigned patdec[kMaxPatterns];

    // tile -> (z tile, y tile, x chunk), z fastest
    const int bz = (int)(L % (unsigned)a.nbz);
    const unsigned r1 = L / (unsigned)a.nbz;
    const int by = (int)(r1 % (unsigned)a.nby);
    const int bx = (int)(r1 / (unsigned)a.nby);
    const int tz = threadIdx.x % 64, ty = threadIdx.x / 64;
    const int n0 = a.s.n0, n1 = a.s.n1, n2 = a.s.n2;
    const int k = bz * TZ + tz * VEC, j = by * TY + ty;  // the lane's first point
    const bool active = k < n2 && j < n1;
    const int kc = min(k, n2 - VEC), jc = min(j, n1 - 1);
    const long plane = (long)n1 * n2;
    const long col = (long)jc * n2 + kc;
    const int zc = tz * VEC + hw;  // tile column of the lane's first point

    // halo sources inside a plane (positions clamped to the mesh; clamped
    // values are never used by a row evaluated here): rows ty < hw and
    // ty >= TY - hw load the VEC values of a row outside the tile; lanes
    // tz < hw and tz >= 64 - hw one value beyond either end of the tile row
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
    } else if (tz >= 64 - hw) {
        const int kk = min(bz * TZ + TZ + tz - (64 - hw), n2 - 1);
        zslot = TZ + hw + tz - (64 - hw);
        zoff = (long)jc * n2 + kk;
    }

    // closure-row tables to LDS (read after the first barrier of the march)
    {
        const int nwords = 3 * kMaxClosureRows * (int)(sizeof(ClosureRow) / sizeof(double));
        const double* src = reinterpret_cast<const double*>(a.rows);
        double* dst = reinterpret_cast<double*>(&crow[0][0]);
        for (int w = threadIdx.x; w < nwords; w += 64 * TY) dst[w] = src[w];
    }
    // pattern decode table: class kPattern0 + p -> table rows (ix, iy, iz),
    // one byte each (read after the first barrier of the march)
    {
        const int ny1 = a.nrows[1] + 1, nz1 = a.nrows[2] + 1;
        for (int p = threadIdx.x; p < kMaxPatterns; p += 64 * TY) {
            const int iz = p % nz1, r = p / nz1;
            patdec[p] = (unsigned)(r / ny1) | ((unsigned)(r % ny1) << 8) | ((unsigned)iz << 16);
        }
    }

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
    // Loads of plane xp into a ring slot.  u0, the accumulator and the class
    // bytes are single-use streams: non-temporal, so that L2 keeps the stage
    // values that neighbouring tiles and chunks re-read.  Stage 0 reads u0
    // as its stage value, so its u0 comes from the queue instead.  Past the
    // chunk every lane loads the plane's first points instead (one line per
    // wave), so that the loads stay unconditional.
    auto load_plane = [&](PlaneLoad<TI, VEC>& ld, int xp) {
        const bool in = xp < x1;
        const long pp = (long)min(xp, x1 - 1) * plane;
        const long cc = in ? col : 0;
        vload<VEC>(us + (long)min(xp + HQ, n0 - 1) * plane + cc, ld.q);
        if (yslot >= 0) {
            vload<VEC>(us + pp + (in ? yoff : 0), ld.y);
        } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e) ld.y[e] = (TI)0;
        }
        ld.z = zslot >= 0 ? us[pp + (in ? zoff : 0)] : (TI)0;
        vload_cls<VEC>(a.cls + pp + cc, ld.c);
        if (STAGE != 0) {
            vload_nt<VEC>(a.u0 + pp + cc, ld.u0);
            vload_nt<VEC>(a.acc + pp + cc, ld.acc);
        } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e) ld.u0[e] = ld.acc[e] = 0.0;
        }
    };
    // x queue: q[m] = us at plane x - HQ + m for m < QW - 1; the last entry
    // (plane x + HQ) is read from the ring slot of plane x
    TI q[QW - 1][VEC];
#pragma unroll
    for (int m = 0; m < QW - 1; ++m) vload<VEC>(us + (long)min(max(x0 - HQ + m, 0), n0 - 1) * plane + col, q[m]);
    // the ring: slot r holds the loads of plane xb + r; they are used in
    // place while that plane is evaluated, then the slot is reloaded with
    // plane xb + r + kMarchAhead, so that no in-flight value is ever copied
    // (the chunk bound is uniform over the workgroup)
    PlaneLoad<TI, VEC> ring[kMarchAhead];
#pragma unroll
    for (int r = 0; r < kMarchAhead; ++r) load_plane(ring[r], x0 + r);

    for (int xb = x0; xb < x1; xb += kMarchAhead)
#pragma unroll
    for (int r = 0; r < kMarchAhead; ++r) {
        const int x = xb + r;
        PlaneLoad<TI, VEC>& ld = ring[r];
        // steps past the chunk (its last group) only keep the ring's loads
        // unconditional
        if (x < x1) {
            const long idx = (long)x * plane + col;
            TI(*t)[LZ] = tile[x & 1];
#pragma unroll
            for (int e = 0; e < VEC; ++e) t[ty + hw][zc + e] = q[HQ][e];
            if (yslot >= 0) {
#pragma unroll
                for (int e = 0; e < VEC; ++e) t[yslot][zc + e] = ld.y[e];
            }
            if (zslot >= 0) t[ty + hw][zslot] = ld.z;
            __syncthreads();
            double vout[VEC], vacc[VEC];
            double sxa[VEC], sya[VEC], sza[VEC];
            unsigned char cl[VEC];
            bool ok[VEC];
            bool all = true;
            // interior circulant rows of every present direction, for every
            // lane (values of lanes that do not use them are discarded): dots
            // from 0.0 in tap order
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const unsigned char c = ld.c[e];
                cl[e] = c;
                ok[e] = active && c != kSpecial;
                all = all && ok[e];
                const int ze = zc + e;
                auto qv = [&](int m) -> TI { return m < QW - 1 ? q[m][e] : ld.q[e]; };
                double sx = 0.0, sy = 0.0, sz = 0.0;
                if constexpr (std::is_same<TI, double>::value) {
#pragma unroll
                    for (int m = 0; m < W; ++m) sx += cx[m] * qv(kXReach + m);
#pragma unroll
                    for (int m = 0; m < W; ++m) sy += cy[m] * t[ty + m][ze];
#pragma unroll
                    for (int m = 0; m < W; ++m) sz += cz[m] * t[ty + hw][ze - hw + m];
                } else {
                    TI vx[W], vy[W], vz[W];
#pragma unroll
                    for (int m = 0; m < W; ++m) {
                        vx[m] = qv(kXReach + m);
                        vy[m] = t[ty + m][ze];
                        vz[m] = t[ty + hw][ze - hw + m];
                    }
                    sx = idot<W>(cx, vx);
                    sy = idot<W>(cy, vy);
                    sz = idot<W>(cz, vz);
                }
                sxa[e] = sx;
                sya[e] = sy;
                sza[e] = sz;
            }
            // closure rows of pattern points replace the interior rows of
            // their directions; each direction's rows of the lane's points in
            // one region (skipped by a wave without such a row).  A row of W
            // columns is summed in a fixed sequence, column by column from
            // 0.0 like the variable-length sum
            unsigned pd[VEC];
#pragma unroll
            for (int e = 0; e < VEC; ++e)
                pd[e] = ok[e] && cl[e] >= kPattern0 ? patdec[cl[e] - kPattern0] : 0u;
            {
                bool any = false;
#pragma unroll
                for (int e = 0; e < VEC; ++e) any = any || (pd[e] >> 16) != 0u;
                if (any) {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const int iz = (int)(pd[e] >> 16);
                        const ClosureRow& cr = crow[2][iz > 0 ? iz - 1 : 0];
                        const TI* c0 = &t[ty + hw][zc + e];
                        double s;
                        if (a.fixw[2]) {
                            // lanes without a z row read the interior frame
                            const int off = iz != 0 ? cr.off : -hw;
                            s = 0.0;
#pragma unroll
                            for (int m = 0; m < W; ++m) s += cr.w[m] * (double)c0[off + m];
                        } else {
                            s = tile_row(cr, c0, 1);
                        }
                        if (iz != 0) sza[e] = s;
                    }
                }
            }
            {
                bool any = false;
#pragma unroll
                for (int e = 0; e < VEC; ++e) any = any || ((pd[e] >> 8) & 0xffu) != 0u;
                if (any) {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const int iy = (int)((pd[e] >> 8) & 0xffu);
                        const ClosureRow& cr = crow[1][iy > 0 ? iy - 1 : 0];
                        const TI* c0 = &t[ty + hw][zc + e];
                        double s;
                        if (a.fixw[1]) {
                            const int off = iy != 0 ? cr.off : -hw;
                            s = 0.0;
#pragma unroll
                            for (int m = 0; m < W; ++m) s += cr.w[m] * (double)c0[(off + m) * LZ];
                        } else {
                            s = tile_row(cr, c0, LZ);
                        }
                        if (iy != 0) sya[e] = s;
                    }
                }
            }
            {
                bool any = false;
#pragma unroll
                for (int e = 0; e < VEC; ++e) any = any || (pd[e] & 0xffu) != 0u;
                if (any) {
#pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const int ix = (int)(pd[e] & 0xffu);
                        if (ix == 0) continue;
                        auto qv = [&](int m) -> TI { return m < QW - 1 ? q[m][e] : ld.q[e]; };
                        // over the queue when the row's columns lie in it,
                        // else from the stage input (the wall planes of a
                        // chunk, uniform over the workgroup); the row's own
                        // columns in column order either way
                        const ClosureRow& cr = crow[0][ix - 1];
                        double s = 0.0;
                        if (cr.off >= -HQ && cr.off + cr.n - 1 <= HQ) {
#pragma unroll
                            for (int m = 0; m < QW; ++m) {
                                const int ci = m - HQ - cr.off;
                                if (ci >= 0 && ci < cr.n) s += cr.w[ci] * (double)qv(m);
                            }
                        } else {
                            const TI* __restrict__ v = us + idx + e + (long)cr.off * plane;
#pragma unroll 1
                            for (int m = 0; m < cr.n; ++m) s += cr.w[m] * (double)v[(long)m * plane];
                        }
                        sxa[e] = s;
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const unsigned char c = cl[e];
                const int ke = k + e;
                // a solid point (kNoRows) has no rows: zero derivatives, no
                // source
                const bool rows = c != kNoRows;
                const double d0 = px && rows ? sxa[e] : 0.0;
                const double d1 = py && rows ? sya[e] : 0.0;
                const double d2 = pz && rows ? sza[e] : 0.0;
                double rhs;
                if (kIsHeat<KIND>) {
                    // laplacian rows added in direction order (plus_eq)
                    double lap = 0.0;
                    if (px) lap += d0;
                    if (py) lap += d1;
                    if (pz) lap += d2;
                    rhs = heat_finish<KIND>(a.s, false, rows, x, j, ke, rows ? lap : 0.0, ss.t_s);
                } else {
                    rhs = wave_combine<KIND>(a.s, false, x, j, ke, d0, d1, d2);
                }
                // stage_update (rk4.cpp:27-46 for one point)
                const double u0v = STAGE == 0 ? (double)q[HQ][e] : ld.u0[e];
                const double aa = (STAGE == 0 ? 0.0 : ld.acc[e]) + ss.c_acc * rhs;
                vacc[e] = aa;
                vout[e] = STAGE < 3 ? u0v + ss.c_next * rhs : u0v + 1.0 * aa;
            }
            // non-temporal stores: the lane's whole vector when the tile
            // evaluated all its points, else point by point (a special point
            // is written by the special work)
            if (all) {
                TO o[VEC];
#pragma unroll
                for (int e = 0; e < VEC; ++e) o[e] = (TO)vout[e];
                if (STAGE < 3) vstore_nt<VEC>(a.acc + idx, vacc);
                vstore_nt<VEC>(out + idx, o);
            } else {
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    if (!ok[e]) continue;
                    if (STAGE < 3) __builtin_nontemporal_store(vacc[e], a.acc + idx + e);
                    __builtin_nontemporal_store((TO)vout[e], out + idx + e);
                }
            }
#pragma unroll
            for (int m = 0; m < QW - 2; ++m)
#pragma unroll
                for (int e = 0; e < VEC; ++e) q[m][e] = q[m + 1][e];
#pragma unroll
            for (int e = 0; e < VEC; ++e) q[QW - 2][e] = ld.q[e];
        }
        load_plane(ld, x + kMarchAhead);
    }
}

// The operands of one row of direction d at a listed point, loaded without
// waiting on each other: the first W columns' values (an interior row's
// taps, a listed row's columns; columns past a row's end repeat its last
// column and are not summed) and the B terms.  `e` is the row's table
// entry.  The weights are read where the row is summed.
template <int W>
struct RowOps {
    double v[W];
    double bw0, bv0, bw1, bv1;
    int n, nb;
    bool row, listed;
};

template <int W, typename TI>
__device__ __forceinline__ void row_ops(RowOps<W>& o, const DevDeriv& d, const SpecialRow& e, unsigned ref,
                                        const TI* __restrict__ u, const double* __restrict__ uR, long idx)
{
    constexpr int hw = W / 2;
    o.listed = ref < kRowInterior;
    o.row = ref != kRowNone;
    o.n = 0;
    o.nb = 0;
    o.bw0 = o.bv0 = o.bw1 = o.bv1 = 0.0;
#pragma unroll
    for (int m = 0; m < W; ++m) o.v[m] = 0.0;
    if (!o.row) return;  // (an absent direction's pointers may be null)
    o.n = o.listed ? e.ncols : W;
    o.nb = o.listed ? e.nb : 0;
    const int off = o.listed ? e.off : -hw;
    const long st = d.stride;
    const int last = o.n - 1;
    if (o.n > 0) {
#pragma unroll
        for (int m = 0; m < W; ++m) o.v[m] = (double)u[idx + (long)(off + min(m, last)) * st];
    }
    if (o.nb > 0) {
        o.bw0 = d.pool[e.bw0];
        o.bv0 = uR[e.bobj0];
    }
    if (o.nb > 1) {
        o.bw1 = d.pool[e.bw1];
        o.bv1 = uR[e.bobj1];
    }
}

// The row from its operands (the same sums as row_at: the dot over the
// row's columns in order from 0.0, B terms after); a listed row longer than
// W continues from memory.
template <int W, typename TI>
__device__ __forceinline__ RowVal row_of(const RowOps<W>& o, const DevDeriv& d, const SpecialRow& e,
                                         const TI* __restrict__ u, long idx)
{
    RowVal r;
    r.row = o.row;
    r.nb = o.nb;
    r.dot = r.b0 = r.b1 = 0.0;
    if (!o.row) return r;
    double dot = 0.0;
    if (o.listed) {
        const double* __restrict__ w = d.pool + e.coef;
        const int last = o.n - 1;
        double wv[W] = {};
        if (o.n > 0) {
#pragma unroll
            for (int m = 0; m < W; ++m) wv[m] = w[min(m, last)];
        }
#pragma unroll
        for (int m = 0; m < W; ++m)
            if (m < o.n) dot += wv[m] * o.v[m];
        if (o.n > W) {
            const TI* __restrict__ v = u + idx + (long)e.off * d.stride;
            for (int c = W; c < o.n; ++c) dot += w[c] * (double)v[(long)c * d.stride];
        }
    } else {
#pragma unroll
        for (int m = 0; m < W; ++m) dot += d.coef[m] * o.v[m];
    }
    r.dot = dot;
    r.b0 = o.bw0 * o.bv0;
    r.b1 = o.bw1 * o.bv1;
    return r;
}

// A free cut point (of a floating or Neumann body): its rows of the three
// derivatives (derivative.cpp:479-485: each row from the running value,
// entries over D then R in stored order; the heat laplacian continues one
// running sum across the directions, the wave's gradient starts each from
// zero), the rhs at the cut point (scalar_wave.cpp:208-238: gG . du with gG
// at the point; heat.cpp:119-162: k lap + source) and the stage update of
// its value and accumulator (rk4.cpp:27-46), all in FP64.  The entries in
// the point's slots load with its record, then their values, before the
// products are summed in order; entries past the slots continue one by one.
// Pointers are picked among values (pick3), never by a run-time array index
// or a choice of addresses.
// One of three pointers by a run-time choice, taken between the three
// values: the values pass through an empty asm statement first, so that the
// compiler cannot turn the choice into one between the addresses of the
// fields that hold them (which would keep an argument structure copied into
// a local variable, as the batch kernel's are, in scratch memory).
template <typename T>
__device__ __forceinline__ T* pick3(int set, T* p0, T* p1, T* p2)
{
    asm volatile("" : "+v"(p0), "+v"(p1), "+v"(p2));
    return set == 0 ? p0 : set == 1 ? p1 : p2;
}

template <int KIND, int STAGE, bool MIXED>
__device__ __forceinline__ void free_item(const SpecialArgs& a, const StageScalars& ss, long f)
{
    using TI = typename StageTypes<MIXED, STAGE>::in;
    const TI* __restrict__ us = static_cast<const TI*>(a.us);
    FreeEnt x8[kFreeSlots];
#pragma unroll
    for (int m = 0; m < kFreeSlots; ++m) x8[m] = a.fents[f * kFreeSlots + m];
    const FreeRec fr = a.frees[f];
    const int set = fr.set;
    const double* __restrict__ uR = pick3(set, a.usR.r[0], a.usR.r[1], a.usR.r[2]);
    const double* pos = pick3(set, a.pos0, a.pos1, a.pos2);
    const double* u0R = pick3(set, a.u0R.r[0], a.u0R.r[1], a.u0R.r[2]);
    const double x = pos[3 * fr.q], y = pos[3 * fr.q + 1], z = pos[3 * fr.q + 2];
    const double u0v = u0R[fr.q];
    const double accv = STAGE == 0 ? 0.0 : a.accF[f];
    const int s1 = fr.n[0], s2 = s1 + fr.n[1], ntot = s2 + fr.n[2];
    double v[kFreeSlots];
#pragma unroll
    for (int m = 0; m < kFreeSlots; ++m) v[m] = x8[m].col >= 0 ? (double)us[x8[m].col] : uR[-1 - x8[m].col];
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;  // heat: r0 is the laplacian's running sum
    // entry e's product goes to the sum of its row (selects on the values,
    // so that the three sums stay in registers)
    auto add = [&](int e, double p) {
        const bool in0 = kIsHeat<KIND> || e < s1;
        const bool in1 = !in0 && e < s2;
        const bool in2 = !in0 && !in1;
        r0 = in0 ? r0 + p : r0;
        r1 = in1 ? r1 + p : r1;
        r2 = in2 ? r2 + p : r2;
    };
#pragma unroll
    for (int m = 0; m < kFreeSlots; ++m)
        if (m < ntot) add(m, x8[m].w * v[m]);
    for (int e = kFreeSlots; e < ntot; ++e) {
        const FreeEnt xe = a.fents[fr.over + e - kFreeSlots];
        add(e, xe.w * (xe.col >= 0 ? (double)us[xe.col] : uR[-1 - xe.col]));
    }
    double rhs;
    if (kIsHeat<KIND>) {
        rhs = r0 * a.s.k;
        rhs += heat_source(a.s, ss.t_s, x, y, z);
    } else {
        rhs = wave_neg_g(a.s, 0, x, y, z) * r0 + wave_neg_g(a.s, 1, x, y, z) * r1 +
              wave_neg_g(a.s, 2, x, y, z) * r2;
    }
    const double aa = accv + ss.c_acc * rhs;
    if (STAGE < 3) a.accF[f] = aa;
    double* outR = pick3(set, a.outR0, a.outR1, a.outR2);
    outR[fr.q] = STAGE < 3 ? u0v + ss.c_next * rhs : u0v + 1.0 * aa;
}

// Point (i, j, k) of face item q: the Dirichlet grid faces in the order
// xmin, xmax, ymin, ymax, zmin, zmax each hold one item per point of their
// plane; a point on several of them belongs to the first (false for the
// items of the others).
__device__ __forceinline__ bool face_point(const DevSys& s, long q, int& i, int& j, int& k)
{
    const int n0 = s.n0, n1 = s.n1, n2 = s.n2;
    bool earlier[6] = {false, false, false, false, false, false};
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        if (s.bc[f] != SHX_BC_DIRICHLET) continue;
        const long size = f < 2 ? (long)n1 * n2 : f < 4 ? (long)n0 * n2 : (long)n0 * n1;
        if (q >= size) {
            q -= size;
            earlier[f] = true;
            continue;
        }
        if (f < 2) {
            i = f == 0 ? 0 : n0 - 1;
            j = (int)(q / n2);
            k = (int)(q % n2);
        } else if (f < 4) {
            j = f == 2 ? 0 : n1 - 1;
            i = (int)(q / n2);
            k = (int)(q % n2);
        } else {
            k = f == 4 ? 0 : n2 - 1;
            i = (int)(q / n1);
            j = (int)(q % n1);
        }
        const bool dup = (earlier[0] && i == 0) || (earlier[1] && i == n0 - 1) || (earlier[2] && j == 0) ||
                         (earlier[3] && j == n1 - 1) || (earlier[4] && k == 0);
        return !dup;
    }
    return false;
}

// Item q of the special work: a free cut point, a listed point, a Dirichlet
// grid-face point, then the cut points of the three directions.  Rows are
// evaluated in FP64 in both forms.  A face point's right-hand side is zero
// and its value exact(t_next); its accumulator is never read (the point's
// value is exact at every stage), so only the value is written.
template <int KIND, int STAGE, int W, bool MIXED, bool FREE = true>
__device__ __forceinline__ void special_item(const SpecialArgs& a, const StageScalars& ss, long q)
{
    using TI = typename StageTypes<MIXED, STAGE>::in;
    using TO = typename StageTypes<MIXED, STAGE>::out;
    const TI* us = static_cast<const TI*>(a.us);
    // the free cut points first: their rows are the longest chains of
    // dependent loads, so they start with the earliest special workgroups
    if constexpr (FREE) {
        if (q < a.nfree) {
            free_item<KIND, STAGE, MIXED>(a, ss, q);
            return;
        }
        q -= a.nfree;
    }
    if (q < a.nspecial) {
        const SpecialRec rec = a.recs[q];
        const long idx = rec.idx;
        int i, j, k;
        point_coords(idx, a.s.n1, a.s.n2, i, j, k);
        const DevOps& ops = a.ops;
        // every load that does not depend on another is issued first: the
        // stage-update operands and the three rows' table entries, then the
        // operands of the three rows together
        const double u0v = a.u0[idx];
        const double accv = STAGE == 0 ? 0.0 : a.acc[idx];
        const unsigned f0 = ops.d[0].present ? rec.ref[0] : kRowNone;
        const unsigned f1 = ops.d[1].present ? rec.ref[1] : kRowNone;
        const unsigned f2 = ops.d[2].present ? rec.ref[2] : kRowNone;
        SpecialRow e0{}, e1{}, e2{};
        if (f0 < kRowInterior) e0 = a.rows0[f0];
        if (f1 < kRowInterior) e1 = a.rows1[f1];
        if (f2 < kRowInterior) e2 = a.rows2[f2];
        RowOps<W> o0, o1, o2;
        row_ops<W>(o0, ops.d[0], e0, f0, us, a.usR.r[0], idx);
        row_ops<W>(o1, ops.d[1], e1, f1, us, a.usR.r[1], idx);
        row_ops<W>(o2, ops.d[2], e2, f2, us, a.usR.r[2], idx);
        const RowVal r0 = row_of<W>(o0, ops.d[0], e0, us, idx);
        const RowVal r1 = row_of<W>(o1, ops.d[1], e1, us, idx);
        const RowVal r2 = row_of<W>(o2, ops.d[2], e2, us, idx);
        double rhs;
        if (kIsHeat<KIND>) {
            double lap = 0.0;
            if (r0.row) lap = row_plus(lap, r0);
            if (r1.row) lap = row_plus(lap, r1);
            if (r2.row) lap = row_plus(lap, r2);
            rhs = heat_finish<KIND>(a.s, on_dirichlet_face(a.s, i, j, k), is_fluid(a.s, idx), i, j, k, lap,
                                    ss.t_s);
        } else {
            double du0 = 0.0, du1 = 0.0, du2 = 0.0;
            if (r0.row) du0 = row_eq(r0);
            if (r1.row) du1 = row_eq(r1);
            if (r2.row) du2 = row_eq(r2);
            rhs = wave_combine<KIND>(a.s, on_dirichlet_face(a.s, i, j, k), i, j, k, du0, du1, du2);
        }
        double v = stage_update<STAGE>(rhs, accv, u0v, a.acc, idx, ss.c_acc, ss.c_next);
        if (on_dirichlet_face(a.s, i, j, k)) v = exact_ijk<KIND>(a.s, i, j, k, ss.t_next);
        static_cast<TO*>(a.out)[idx] = (TO)v;
        return;
    }
    q -= a.nspecial;
    if (q < a.nface) {
        int i, j, k;
        if (face_point(a.s, q, i, j, k))
            static_cast<TO*>(a.out)[((long)i * a.s.n1 + j) * a.s.n2 + k] =
                (TO)exact_ijk<KIND>(a.s, i, j, k, ss.t_next);
        return;
    }
    q -= a.nface;
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
        return;
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
template <int KIND, int STAGE, int W, int TY, int VEC, bool MIXED, bool FREE = true>
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
        special_item<KIND, STAGE, W, MIXED, FREE>(sp, ss, blk * (64 * TY) + threadIdx.x);
        return;
    }
    const unsigned m = mi * 8u + lane8;
    if (m >= a.nblocks) return;
    march_tile<KIND, STAGE, W, TY, VEC, MIXED>(a, ss, xcd_remap(m, a.nblocks));
}

template <int KIND, int STAGE, int W, int TY, int VEC, bool MIXED>
__global__ __launch_bounds__(64 * TY) void k_stage(MarchArgs a, SpecialArgs sp, StageScalars ss,
                                                   unsigned sg, unsigned mg, int mode)
{
    stage_block<KIND, STAGE, W, TY, VEC, MIXED>(a, sp, ss, blockIdx.x, sg, mg, mode);
}

// The same stage for the heat systems, held to four waves per SIMD: their
// manufactured solutions' out-of-line evaluations (sines, cosines and
// exponentials at cut points, reached only by special items) would
// otherwise set the register count of the whole kernel; under the bound
// they spill instead, where they run.
template <int KIND, int STAGE, int W, int TY, int VEC, bool MIXED>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(4))) void k_stage_heat(
    MarchArgs a, SpecialArgs sp, StageScalars ss, unsigned sg, unsigned mg, int mode)
{
    stage_block<KIND, STAGE, W, TY, VEC, MIXED>(a, sp, ss, blockIdx.x, sg, mg, mode);
}

// RK4 stage constants (rk4.cpp:13-14), as the host's
constexpr double kStageTime[4] = {0.0, 0.5, 0.5, 1.0};
constexpr double kStageWeight[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

// One fused RK4 stage of a batch of problems in one launch: problem p owns
// workgroups [start[p], start[p+1]) (multiples of 8, so that tile workgroups
// keep their XCD); its arguments and clock live in device memory, so that a
// captured launch stays valid at any time.  The stage scalars are formed as
// the single-problem host path forms them (t = t0 + n dt).
template <int KIND, int STAGE, int W, int TY, int VEC, bool MIXED>
__device__ __forceinline__ void batch_block(const BatchEntry* __restrict__ ent, const BatchClock* __restrict__ clk,
                                            const unsigned* __restrict__ start, int nprob)
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
    // problems with free cut points (floating bodies) carry them in the
    // special work, as the single-problem stage does
    stage_block<KIND, STAGE, W, TY, VEC, MIXED, true>(m, sp, ss, blockIdx.x - start[lo], sg, mg, mode);
}

// Held to four waves per SIMD, as k_stage_heat: the batch kernel carries
// the special work of every system family it serves (the FP64 wave form
// needs 133-135 registers unbounded, 111-127 and no scratch bounded).
template <int KIND, int STAGE, int W, int TY, int VEC, bool MIXED>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_stage(
    const BatchEntry* __restrict__ ent, const BatchClock* __restrict__ clk, const unsigned* __restrict__ start,
    int nprob)
{
    batch_block<KIND, STAGE, W, TY, VEC, MIXED>(ent, clk, start, nprob);
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
#define SHX_STAGE_CASE(TY, VEC)                                                                      \
    case TY * 10 + VEC:                                                                              \
        if constexpr (kIsHeat<KIND>)                                                                 \
            k_stage_heat<KIND, STAGE, W, TY, VEC, MIXED><<<grid, 64 * TY, 0, st>>>(a, sp, ss, s_groups, m_groups, md); \
        else                                                                                         \
            k_stage<KIND, STAGE, W, TY, VEC, MIXED><<<grid, 64 * TY, 0, st>>>(a, sp, ss, s_groups, m_groups, md); \
        return 0;
        switch (ty * 10 + a.vec) {
            SHX_STAGE_CASE(4, 1)
            SHX_STAGE_CASE(8, 1)
            SHX_STAGE_CASE(4, 2)
            SHX_STAGE_CASE(8, 2)
        default: return -1;
        }
#undef SHX_STAGE_CASE
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
    const long nitems = sp.nspecial + sp.nface + sp.nR0 + sp.nR1 + sp.nR2 + sp.nfree;
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
#define SHX_BATCH_CASE(TY, VEC)                                                                      \
    case TY * 10 + VEC:                                                                              \
        k_batch_stage<KIND, STAGE, W, TY, VEC, MIXED><<<bl.grid, 64 * TY, 0, st>>>(bl.ent, bl.clk, bl.start, \
                                                                                 bl.nprob);          \
        return 0;
    switch (ty * 10 + bl.vec) {
        SHX_BATCH_CASE(4, 1)
        SHX_BATCH_CASE(8, 1)
        SHX_BATCH_CASE(4, 2)
        SHX_BATCH_CASE(8, 2)
    default: return -1;
    }
#undef SHX_BATCH_CASE
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
