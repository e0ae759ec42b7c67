// shx_march.hip -- fused RK4 stage kernels that march along x.
//
// A workgroup owns a (y, z) tile of TY x 64 points and walks a chunk of x
// planes.  Each thread keeps the five x-neighbours of its point in registers
// (a queue shifted by one plane per step), the current plane is staged in LDS
// with two-point y and z halos, so every stage value is fetched from HBM about
// once.  Rows whose position is not in one of the line's interior intervals
// (closures, B terms, solid points) take the segment-table slow path, which
// reads global memory.  Arithmetic is the same as the point-wise kernel of
// shx_systems.hip, term for term, so results are bit-identical.
#include <hip/hip_runtime.h>

#include "shx_sys.hpp"

namespace shx {

namespace {

__device__ __forceinline__ bool in_fast(unsigned long long fd, int q)
{
    const int lo0 = (int)(fd & 0xffffu), hi0 = (int)((fd >> 16) & 0xffffu);
    const int lo1 = (int)((fd >> 32) & 0xffffu), hi1 = (int)(fd >> 48);
    return (q >= lo0 && q < hi0) || (q >= lo1 && q < hi1);
}

// interior 5-point dot from 0.0, left to right
__device__ __forceinline__ double dot5(const double* c, double a, double b, double m, double d,
                                       double e)
{
    double s = 0.0;
    s += c[0] * a;
    s += c[1] * b;
    s += c[2] * m;
    s += c[3] * d;
    s += c[4] * e;
    return s;
}

template <int STAGE, int TY>
__global__ __launch_bounds__(64 * TY) void k_wave_stage_march(MarchArgs a)
{
    constexpr int TZ = 64;
    __shared__ double tile[TY + 4][TZ + 4];

    // workgroup -> (z tile, y tile, x chunk), contiguous per XCD, z fastest
    const unsigned L = xcd_remap(blockIdx.x, a.nblocks);
    const int bz = (int)(L % (unsigned)a.nbz);
    const unsigned r1 = L / (unsigned)a.nbz;
    const int by = (int)(r1 % (unsigned)a.nby);
    const int bx = (int)(r1 / (unsigned)a.nby);
    const int tz = threadIdx.x % TZ, ty = threadIdx.x / TZ;
    const int n0 = a.s.n0, n1 = a.s.n1, n2 = a.s.n2;
    const int k = bz * TZ + tz, j = by * TY + ty;
    const bool active = k < n2 && j < n1;
    const int x0 = bx * a.xchunk;
    const int x1 = min(x0 + a.xchunk, n0);
    const long plane = (long)n1 * n2;
    const long col = (long)j * n2 + k;  // offset of (j, k) inside a plane

    const DevDeriv& dx = a.ops.d[0];
    const DevDeriv& dy = a.ops.d[1];
    const DevDeriv& dz = a.ops.d[2];
    double cx[5], cy[5], cz[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        cx[q] = dx.present ? dx.coef[q] : 0.0;
        cy[q] = dy.present ? dy.coef[q] : 0.0;
        cz[q] = dz.present ? dz.coef[q] : 0.0;
    }
    const unsigned long long fx = (active && dx.present) ? dx.fast[j * n2 + k] : 0ull;

    const double* __restrict__ us = a.us;
    auto ld = [&](int x) -> double { return (active && x >= 0 && x < n0) ? us[(long)x * plane + col] : 0.0; };
    double qm2 = ld(x0 - 2), qm1 = ld(x0 - 1), q0 = ld(x0), qp1 = ld(x0 + 1), qp2;

    for (int x = x0; x < x1; ++x) {
        // issue every global load of this plane before the barrier: the queue
        // head, the LDS halos, the y/z line descriptors and u0/acc at the point
        qp2 = ld(x + 2);
        const long pbase = (long)x * plane;
        double hy = 0.0, hz = 0.0;
        int hy_row = -1, hz_col = -1;
        if (ty < 2) {
            const int jj = by * TY - 2 + ty;
            hy_row = ty;
            if (k < n2 && jj >= 0 && jj < n1) hy = us[pbase + (long)jj * n2 + k];
        } else if (ty >= TY - 2) {
            const int jj = by * TY + TY + (ty - (TY - 2));
            hy_row = ty + 4;
            if (k < n2 && jj < n1) hy = us[pbase + (long)jj * n2 + k];
        }
        double hy2 = 0.0;
        int hy2_row = -1;
        if (TY == 2) {  // two rows of threads also load the bottom halo
            const int jj = by * TY + TY + ty;
            hy2_row = ty + 4;
            if (k < n2 && jj < n1) hy2 = us[pbase + (long)jj * n2 + k];
        }
        if (tz < 2) {
            const int kk = bz * TZ - 2 + tz;
            hz_col = tz;
            if (j < n1 && kk >= 0 && kk < n2) hz = us[pbase + (long)j * n2 + kk];
        } else if (tz >= TZ - 2) {
            const int kk = bz * TZ + TZ + (tz - (TZ - 2));
            hz_col = tz + 4;
            if (j < n1 && kk < n2) hz = us[pbase + (long)j * n2 + kk];
        }
        const long idx = pbase + col;
        const unsigned long long fy = (active && dy.present) ? dy.fast[x * n2 + k] : 0ull;
        const unsigned long long fz = (active && dz.present) ? dz.fast[x * n1 + j] : 0ull;
        const double u0v = active ? a.u0[idx] : 0.0;
        const double accv = (active && STAGE > 0) ? a.acc[idx] : 0.0;

        tile[ty + 2][tz + 2] = q0;
        if (hy_row >= 0) tile[hy_row][tz + 2] = hy;
        if (hy2_row >= 0) tile[hy2_row][tz + 2] = hy2;
        if (hz_col >= 0) tile[ty + 2][hz_col] = hz;
        __syncthreads();

        if (active) {
            double du[3] = {0.0, 0.0, 0.0};
            // x: queue
            if (dx.present) {
                if (in_fast(fx, x)) {
                    du[0] = dot5(cx, qm2, qm1, q0, qp1, qp2);
                } else {
                    const RowVal r = row_slow<true>(dx.pair_ptr, dx.blocks, dx.pool, dx.coef, dx.width,
                                                    dx.stride, us, a.usR[0], idx, j * n2 + k, x);
                    if (r.row) du[0] = row_eq(r);
                }
            }
            // y: LDS column
            if (dy.present) {
                if (in_fast(fy, j)) {
                    du[1] = dot5(cy, tile[ty][tz + 2], tile[ty + 1][tz + 2], tile[ty + 2][tz + 2],
                                 tile[ty + 3][tz + 2], tile[ty + 4][tz + 2]);
                } else {
                    const RowVal r = row_slow<true>(dy.pair_ptr, dy.blocks, dy.pool, dy.coef, dy.width,
                                                    dy.stride, us, a.usR[1], idx, x * n2 + k, j);
                    if (r.row) du[1] = row_eq(r);
                }
            }
            // z: LDS row
            if (dz.present) {
                if (in_fast(fz, k)) {
                    du[2] = dot5(cz, tile[ty + 2][tz], tile[ty + 2][tz + 1], tile[ty + 2][tz + 2],
                                 tile[ty + 2][tz + 3], tile[ty + 2][tz + 4]);
                } else {
                    const RowVal r = row_slow<true>(dz.pair_ptr, dz.blocks, dz.pool, dz.coef, dz.width,
                                                    dz.stride, us, a.usR[2], idx, x * n1 + j, k);
                    if (r.row) du[2] = row_eq(r);
                }
            }
            double g0 = 0.0, g1 = 0.0, g2 = 0.0;
            const bool face = on_dirichlet_face(a.s, x, j, k);
            if (!face) {
                const double px = a.s.x[x], py = a.s.y[j], pz = a.s.z[k];
                g0 = wave_neg_g(a.s, 0, px, py, pz);
                g1 = wave_neg_g(a.s, 1, px, py, pz);
                g2 = wave_neg_g(a.s, 2, px, py, pz);
            }
            const double rhs = g0 * du[0] + g1 * du[1] + g2 * du[2];
            const double acc = (STAGE == 0 ? 0.0 : accv) + a.c_acc * rhs;
            double v;
            if (STAGE < 3) {
                a.acc[idx] = acc;
                v = u0v + a.c_next * rhs;
            } else {
                v = u0v + 1.0 * acc;
            }
            if (face) v = wave_exact(a.s, a.s.x[x], a.s.y[j], a.s.z[k], a.t_next);
            a.out[idx] = v;
        }
        __syncthreads();
        qm2 = qm1;
        qm1 = q0;
        q0 = qp1;
        qp1 = qp2;
    }
}

}  // namespace

int launch_wave_stage_march(hipStream_t st, int stage, int ty, const MarchArgs& a)
{
#define SHX_MARCH_CASE(S, T) \
    if (stage == S && ty == T) { k_wave_stage_march<S, T><<<a.nblocks, 64 * T, 0, st>>>(a); return 0; }
    SHX_MARCH_CASE(0, 4) SHX_MARCH_CASE(1, 4) SHX_MARCH_CASE(2, 4) SHX_MARCH_CASE(3, 4)
    SHX_MARCH_CASE(0, 8) SHX_MARCH_CASE(1, 8) SHX_MARCH_CASE(2, 8) SHX_MARCH_CASE(3, 8)
    SHX_MARCH_CASE(0, 2) SHX_MARCH_CASE(1, 2) SHX_MARCH_CASE(2, 2) SHX_MARCH_CASE(3, 2)
#undef SHX_MARCH_CASE
    return -1;
}

}  // namespace shx
