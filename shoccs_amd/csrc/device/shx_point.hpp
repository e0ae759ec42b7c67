// shx_point.hpp -- point-wise pieces of the right-hand sides and of one fused
// RK4 stage, shared by every kernel that evaluates them (the point-wise tiles,
// the special-point list of the x-marching form, the reference schedule).
// One definition of each formula keeps the kernels bit-identical.
#pragma once

#include <hip/hip_runtime.h>

#include "shx_internal.hpp"
#include "shx_sys.hpp"

namespace shx {

// gauss3d exponent (src/mms/gauss3d.cpp:14-20)
__device__ __forceinline__ double g3_expo(const double* g, double x, double y, double z)
{
    const double *c = g, *v = g + 3;
    return (-0.5 * pow((x - c[0]), 2) * pow(v[0], -2) +
            (-0.5 * pow((y - c[1]), 2) * pow(v[1], -2) + -0.5 * pow((z - c[2]), 2) * pow(v[2], -2)));
}

// The gaussian exponent of the mesh's dimension (build_ms_gauss,
// src/mms/mms.cpp:16-31: gauss1d.cpp:14-16, gauss2d.cpp:14-19, gauss3d).
__device__ __forceinline__ double g_expo(int dims, const double* g, double x, double y, double z)
{
    const double *c = g, *v = g + 3;
    if (dims == 1) return -0.5 * pow((x - c[0]), 2) * pow(v[0], -2);
    if (dims == 2) return (-0.5 * pow((x - c[0]), 2) * pow(v[0], -2) + -0.5 * pow((y - c[1]), 2) * pow(v[1], -2));
    return g3_expo(g, x, y, z);
}

// Manufactured solutions; scalar arguments only (called out of line).
static __device__ __noinline__ double mms_call(int kind, int dims, int n,
                                               const double* __restrict__ p, double t, double x,
                                               double y, double z)
{
    if (kind == SHX_MMS_SINPROD) {
        double v = sin(kTwoPi * x) * sin(kTwoPi * y);
        if (dims == 3) v = v * sin(kTwoPi * z);
        return v * exp(-t);
    }
    double sol = 0;
    for (int i = 0; i < n; ++i) {
        const double* g = p + 8 * i;
        sol += exp(g_expo(dims, g, x, y, z)) * g[6] * cos(t * g[7]);
    }
    return sol;
}

static __device__ __noinline__ double mms_ddt(int kind, int dims, int n,
                                              const double* __restrict__ p, double t, double x,
                                              double y, double z)
{
    if (kind == SHX_MMS_SINPROD) return -mms_call(kind, dims, n, p, t, x, y, z);
    double sol = 0;
    for (int i = 0; i < n; ++i) {
        const double* g = p + 8 * i;
        sol += -exp(g_expo(dims, g, x, y, z)) * g[6] * g[7] * sin(t * g[7]);
    }
    return sol;
}

static __device__ __noinline__ double mms_lap(int kind, int dims, int n,
                                              const double* __restrict__ p, double t, double x,
                                              double y, double z)
{
    if (kind == SHX_MMS_SINPROD) return -(dims * kTwoPi * kTwoPi) * mms_call(kind, dims, n, p, t, x, y, z);
    double sol = 0;
    if (dims == 1) {  // gauss1d.cpp:47-60
        for (int i = 0; i < n; ++i) {
            const double* c = p + 8 * i;
            const double* v = c + 3;
            sol += c[6] * cos(t * c[7]) *
                   (exp(-0.5 * pow((x - c[0]), 2) * pow(v[0], -2)) * pow((x - c[0]), 2) * pow(v[0], -4) -
                    exp(-0.5 * pow((x - c[0]), 2) * pow(v[0], -2)) * pow(v[0], -2));
        }
        return sol;
    }
    if (dims == 2) {  // gauss2d.cpp:65-85
        for (int i = 0; i < n; ++i) {
            const double* c = p + 8 * i;
            const double* v = c + 3;
            sol += exp((-0.5 * pow((x - c[0]), 2) * pow(v[0], -2) + -0.5 * pow((y - c[1]), 2) * pow(v[1], -2))) *
                   c[6] * cos(t * c[7]) * pow(v[0], -4) * pow(v[1], -4) *
                   (((pow(y, 2) * pow(v[0], 4) +
                      (-2 * y * c[1] * pow(v[0], 4) +
                       (pow(c[1], 2) * pow(v[0], 4) +
                        (pow(x, 2) * pow(v[1], 4) +
                         (-2 * x * c[0] * pow(v[1], 4) + pow(c[0], 2) * pow(v[1], 4)))))) -
                     pow(v[0], 2) * pow(v[1], 4)) -
                    pow(v[0], 4) * pow(v[1], 2));
        }
        return sol;
    }
    for (int i = 0; i < n; ++i) {
        const double* c = p + 8 * i;
        const double* v = c + 3;
        sol += exp(0.5 * ((-1 * pow((z - c[2]), 2) * pow(v[2], -2) -
                           pow((y - c[1]), 2) * pow(v[1], -2)) -
                          pow((x - c[0]), 2) * pow(v[0], -2))) *
               c[6] * cos(t * c[7]) * pow(v[0], -4) * pow(v[1], -4) * pow(v[2], -4) *
               ((((pow(z, 2) * pow(v[0], 4) * pow(v[1], 4) +
                   (-2 * z * c[2] * pow(v[0], 4) * pow(v[1], 4) +
                    (pow(c[2], 2) * pow(v[0], 4) * pow(v[1], 4) +
                     (pow(y, 2) * pow(v[0], 4) * pow(v[2], 4) +
                      (-2 * y * c[1] * pow(v[0], 4) * pow(v[2], 4) +
                       (pow(c[1], 2) * pow(v[0], 4) * pow(v[2], 4) +
                        (pow(x, 2) * pow(v[1], 4) * pow(v[2], 4) +
                         (-2 * x * c[0] * pow(v[1], 4) * pow(v[2], 4) +
                          pow(c[0], 2) * pow(v[1], 4) * pow(v[2], 4))))))))) -
                  pow(v[0], 2) * pow(v[1], 4) * pow(v[2], 4)) -
                 pow(v[0], 4) * pow(v[1], 2) * pow(v[2], 4)) -
                pow(v[0], 4) * pow(v[1], 4) * pow(v[2], 2));
    }
    return sol;
}

// Gradient component d of the manufactured solution (the Neumann face data of
// heat.cpp:318-346): the sin product's closed form, and the gaussians'
// gradient (src/mms/gauss3d.cpp:45-75, gauss2d.cpp:37-57, gauss1d.cpp:30-40),
// each term in the reference's order.
static __device__ __noinline__ double mms_grad(int kind, int dims, int n, const double* __restrict__ p,
                                               double t, double x, double y, double z, int d)
{
    if (kind == SHX_MMS_SINPROD) {
        const double sx = sin(kTwoPi * x), sy = sin(kTwoPi * y);
        const double cx = cos(kTwoPi * x), cy = cos(kTwoPi * y);
        const double sz = dims == 3 ? sin(kTwoPi * z) : 1.0;
        const double cz = dims == 3 ? cos(kTwoPi * z) : 0.0;
        const double e = exp(-t);
        if (d == 0) return kTwoPi * cx * sy * sz * e;
        if (d == 1) return kTwoPi * sx * cy * sz * e;
        return dims == 3 ? kTwoPi * sx * sy * cz * e : 0.0;
    }
    if (d >= dims) return 0.0;  // gauss1d/gauss2d leave the other components zero
    const double loc[3] = {x, y, z};
    double sol = 0;
    for (int i = 0; i < n; ++i) {
        const double* g = p + 8 * i;
        const double base = -exp(g_expo(dims, g, x, y, z)) * g[6] * cos(t * g[7]);
        sol += base * (loc[d] - g[d]) * pow(g[3 + d], -2);
    }
    return sol;
}

// heat source S_t - k lap S (heat.cpp:111-117)
__device__ __forceinline__ double heat_source(const DevSys& s, double t, double x, double y, double z)
{
    return mms_ddt(s.mms_kind, s.mms_dims, s.mms_n, s.mms, t, x, y, z) -
           s.k * mms_lap(s.mms_kind, s.mms_dims, s.mms_n, s.mms, t, x, y, z);
}

__device__ __forceinline__ double exact_at(const DevSys& s, double x, double y, double z, double t)
{
    return s.kind == SHX_SYS_HEAT ? mms_call(s.mms_kind, s.mms_dims, s.mms_n, s.mms, t, x, y, z)
                                  : wave_exact(s, x, y, z, t);
}

// Exact solution with the system family fixed at compile time, so that the
// scalar-wave kernels carry no call into the manufactured solutions.
template <int KIND>
__device__ __forceinline__ double exact_of(const DevSys& s, double x, double y, double z, double t)
{
    // heat: the manufactured solution's own (out-of-line) evaluation
    if (KIND == SHX_SYS_HEAT || KIND == kHeatGauss) return mms_call(s.mms_kind, s.mms_dims, s.mms_n, s.mms, t, x, y, z);
    return wave_exact(s, x, y, z, t);
}

// Sin-product MMS at mesh point (i, j, k) from the per-axis tables: the same
// operations, in the same order, as mms_call on the point's coordinates.
__device__ __forceinline__ double sinprod_at(const DevSys& s, int i, int j, int k, double t)
{
    double v = s.sinx[i] * s.siny[j];
    if (s.mms_dims == 3) v = v * s.sinz[k];
    return v * exp(-t);
}

// Exact solution at mesh point (i, j, k).
template <int KIND>
__device__ __forceinline__ double exact_ijk(const DevSys& s, int i, int j, int k, double t)
{
    if (KIND == SHX_SYS_HEAT) return sinprod_at(s, i, j, k, t);
    return exact_of<KIND>(s, s.x[i], s.y[j], s.z[k], t);
}

// Heat source S_t - k lap S at mesh point (i, j, k) (heat.cpp:111-117); the
// sin-product terms as mms_ddt and mms_lap form them.
template <int KIND>
__device__ __forceinline__ double heat_source_ijk(const DevSys& s, double t, int i, int j, int k)
{
    if (KIND == SHX_SYS_HEAT) {
        const double v = sinprod_at(s, i, j, k, t);
        return -v - s.k * (-(s.mms_dims * kTwoPi * kTwoPi) * v);
    }
    return heat_source(s, t, s.x[i], s.y[j], s.z[k]);
}

// Scalar-wave combination gG . du at one point of D (scalar_wave.cpp:208-238:
// left to right); gG is zero on Dirichlet grid faces (scalar_wave.cpp:70-95).
// KIND is the wave family (constant velocity: gG = -a needs no coordinates).
template <int KIND>
__device__ __forceinline__ double wave_combine(const DevSys& s, bool face, int i, int j, int k,
                                               double du0, double du1, double du2)
{
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (!face) {
        if (KIND == SHX_SYS_WAVE_LINEAR) {
            g0 = -s.vel[0];
            g1 = -s.vel[1];
            g2 = -s.vel[2];
        } else {
            const double x = s.x[i], y = s.y[j], z = s.z[k];
            g0 = wave_neg_g(s, 0, x, y, z);
            g1 = wave_neg_g(s, 1, x, y, z);
            g2 = wave_neg_g(s, 2, x, y, z);
        }
    }
    return g0 * du0 + g1 * du1 + g2 * du2;
}

// Heat right-hand side from the accumulated laplacian rows (heat.cpp:96-128):
// scale by the diffusivity, zero on Dirichlet faces, add the source on fluid.
template <int KIND>
__device__ __forceinline__ double heat_finish(const DevSys& s, bool face, bool fluid, int i, int j,
                                              int k, double lap, double t)
{
    double acc = lap * s.k;
    if (face) return 0.0;
    if (fluid) acc += heat_source_ijk<KIND>(s, t, i, j, k);
    return acc;
}

// Right-hand side at one point of D (interior stencil width W); KIND is the
// system family of the problem (SHX_SYS_*).
template <int KIND, int W>
__device__ __forceinline__ double rhs_point(const DevSys& s, const DevOps& ops,
                                            const double* __restrict__ u, const RPtrs& uR,
                                            long idx, int i, int j, int k, double t)
{
    if (kIsHeat<KIND>) {
        // laplacian: du = 0; du += O_x u; du += B_x u_Rx; then y, z (laplacian.cpp:37-46)
        double acc = 0.0;
        if (ops.d[0].present) {
            const RowVal r = row_at<0, W, true>(ops.d[0], u, uR.r[0], idx, i, j, k, s.n1, s.n2);
            if (r.row) acc = row_plus(acc, r);
        }
        if (ops.d[1].present) {
            const RowVal r = row_at<1, W, true>(ops.d[1], u, uR.r[1], idx, i, j, k, s.n1, s.n2);
            if (r.row) acc = row_plus(acc, r);
        }
        if (ops.d[2].present) {
            const RowVal r = row_at<2, W, true>(ops.d[2], u, uR.r[2], idx, i, j, k, s.n1, s.n2);
            if (r.row) acc = row_plus(acc, r);
        }
        return heat_finish<KIND>(s, on_dirichlet_face(s, i, j, k), is_fluid(s, idx), i, j, k, acc, t);
    } else {
        double du0 = 0.0, du1 = 0.0, du2 = 0.0;
        if (ops.d[0].present) {
            const RowVal r = row_at<0, W, true>(ops.d[0], u, uR.r[0], idx, i, j, k, s.n1, s.n2);
            if (r.row) du0 = row_eq(r);
        }
        if (ops.d[1].present) {
            const RowVal r = row_at<1, W, true>(ops.d[1], u, uR.r[1], idx, i, j, k, s.n1, s.n2);
            if (r.row) du1 = row_eq(r);
        }
        if (ops.d[2].present) {
            const RowVal r = row_at<2, W, true>(ops.d[2], u, uR.r[2], idx, i, j, k, s.n1, s.n2);
            if (r.row) du2 = row_eq(r);
        }
        return wave_combine<KIND>(s, on_dirichlet_face(s, i, j, k), i, j, k, du0, du1, du2);
    }
}

// Stage update of one point given its right-hand side (rk4.cpp:27-46 for one
// point, over three registers):
//   acc' = acc + c_acc * rhs  (acc = 0 on stage 0)
//   stages 0-2: out = u0 + c_next * rhs;  stage 3: out = u0 + 1.0 * acc'
// Returns the value for `out`; stores acc' for stages 0-2.
template <int STAGE>
__device__ __forceinline__ double stage_update(double rhs, double accv, double u0v,
                                               double* __restrict__ acc, long idx, double c_acc,
                                               double c_next)
{
    const double a = (STAGE == 0 ? 0.0 : accv) + c_acc * rhs;
    if (STAGE < 3) {
        acc[idx] = a;
        return u0v + c_next * rhs;
    }
    return u0v + 1.0 * a;
}

// Stage update and store of one point from its right-hand side; Dirichlet
// faces take exact(t_next).  The output register may be FP32 (mixed form).
template <int KIND, int STAGE, typename TO>
__device__ __forceinline__ void stage_store(const DevSys& s, double rhs, const double* u0,
                                            double* __restrict__ acc, TO* out, long idx, int i,
                                            int j, int k, double c_acc, double c_next,
                                            double t_next)
{
    double v = stage_update<STAGE>(rhs, STAGE == 0 ? 0.0 : acc[idx], u0[idx], acc, idx, c_acc,
                                   c_next);
    if (on_dirichlet_face(s, i, j, k)) v = exact_ijk<KIND>(s, i, j, k, t_next);
    out[idx] = (TO)v;
}

// One point of one fused RK4 stage.
template <int KIND, int STAGE, int W>
__device__ __forceinline__ void stage_point(const DevSys& s, const DevOps& ops,
                                            const double* __restrict__ us, const RPtrs& usR,
                                            const double* u0, double* __restrict__ acc,
                                            double* out, long idx, int i, int j, int k,
                                            double t_s, double c_acc, double c_next,
                                            double t_next)
{
    const double rhs = rhs_point<KIND, W>(s, ops, us, usR, idx, i, j, k, t_s);
    stage_store<KIND, STAGE>(s, rhs, u0, acc, out, idx, i, j, k, c_acc, c_next, t_next);
}

}  // namespace shx
