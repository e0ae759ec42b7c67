// shx_systems.hip -- scalar-wave and heat right-hand sides, boundary updates,
// statistics and the RK4 step (reference slot schedule and fused form) of the
// SHOCCS device library (gfx950).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "shx_internal.hpp"
#include "shx_state.hpp"
#include "shx_point.hpp"
#include "shx_sys.hpp"
#include "shx_system_state.hpp"

using namespace shx;


// --------------------------------------------------------------------- kernels
namespace {

template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_rhs(Tiling t, DevSys s, DevOps ops,
                                                const double* __restrict__ u, RPtrs uR,
                                                double* __restrict__ rhs, double time)
{
    int i, j, k;
    long idx;
    if (!tile_point(t, i, j, k, idx)) return;
    rhs[idx] = rhs_point<KIND, W>(s, ops, u, uR, idx, i, j, k, time);
}

// One fused RK4 stage over three registers, point-wise tiles (stage_point).
template <int KIND, int STAGE, int W>
__global__ __launch_bounds__(kBlock) void k_rk4_stage(Tiling t, DevSys s, DevOps ops,
                                                      const double* __restrict__ us, RPtrs usR,
                                                      const double* u0, double* __restrict__ acc,
                                                      double* out, double t_s, double c_acc,
                                                      double c_next, double t_next)
{
    int i, j, k;
    long idx;
    if (!tile_point(t, i, j, k, idx)) return;
    stage_point<KIND, STAGE, W>(s, ops, us, usR, u0, acc, out, idx, i, j, k, t_s, c_acc, c_next,
                                t_next);
}

// Dirichlet cut points take the exact solution (update_boundary on R).
__global__ void k_r_exact(DevSys s, long n, const double* __restrict__ pos,
                          const signed char* __restrict__ dirich, double* __restrict__ uR,
                          double t)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !dirich[i]) return;
    uR[i] = exact_at(s, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], t);
}

// Dirichlet grid faces of D take the exact solution.
__global__ void k_face_exact(DevSys s, long n, const int* __restrict__ faces,
                             double* __restrict__ u, double t)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long idx = faces[q];
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    u[idx] = exact_at(s, s.x[i], s.y[j], s.z[k], t);
}

__global__ void k_exact_D(DevSys s, long N, double* __restrict__ out, double t)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    out[idx] = exact_at(s, s.x[i], s.y[j], s.z[k], t);
}

__global__ void k_exact_R(DevSys s, long n, const double* __restrict__ pos, double* __restrict__ out,
                          double t)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = exact_at(s, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], t);
}

// sin(2 pi x) per mesh coordinate, for the sin-product MMS
__global__ void k_sin_table(const double* __restrict__ x, long n, double* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = sin(kTwoPi * x[i]);
}

// initialize_scalar_field on the fluid gather
__global__ void k_init_fluid(DevSys s, long n, const int* __restrict__ fluid, double* __restrict__ u,
                             double t)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long idx = fluid[q];
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    u[idx] = exact_at(s, s.x[i], s.y[j], s.z[k], t);
}

// gG fields of the reference (scalar_wave.cpp:70-95), stored for the unfused form
__global__ void k_gG_D(DevSys s, long N, double* __restrict__ gx, double* __restrict__ gy,
                       double* __restrict__ gz)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    const double x = s.x[i], y = s.y[j], z = s.z[k];
    const bool f = on_dirichlet_face(s, i, j, k);
    gx[idx] = f ? 0.0 : wave_neg_g(s, 0, x, y, z);
    gy[idx] = f ? 0.0 : wave_neg_g(s, 1, x, y, z);
    gz[idx] = f ? 0.0 : wave_neg_g(s, 2, x, y, z);
}

// scalar_wave.cpp:70-95 on the cut points: -g at those of non-Dirichlet
// bodies, zero at Dirichlet ones
__global__ void k_gG_R(DevSys s, long n, const double* __restrict__ pos, const signed char* __restrict__ dirich,
                       double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    const bool f = dirich[i] != 0;
    gx[i] = f ? 0.0 : wave_neg_g(s, 0, x, y, z);
    gy[i] = f ? 0.0 : wave_neg_g(s, 1, x, y, z);
    gz[i] = f ? 0.0 : wave_neg_g(s, 2, x, y, z);
}

// scalar_wave.cpp:294-317, one buffer
__global__ void k_dot3(long n, const double* __restrict__ gx, const double* __restrict__ dx,
                       const double* __restrict__ gy, const double* __restrict__ dy,
                       const double* __restrict__ gz, const double* __restrict__ dz,
                       double* __restrict__ o)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    o[i] = gx[i] * dx[i] + gy[i] * dy[i] + gz[i] * dz[i];
}

__global__ void k_scale(long n, double* __restrict__ a, double c)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] *= c;
}

// heat fill_source into a stored buffer (heat.cpp:111-117)
__global__ void k_src_D(DevSys s, long N, double* __restrict__ src, double t)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    src[idx] = heat_source(s, t, s.x[i], s.y[j], s.z[k]);
}

__global__ void k_gather_add(long n, const int* __restrict__ sel, const double* __restrict__ src,
                             double* __restrict__ dst)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long idx = sel[q];
    dst[idx] += src[idx];
}

__global__ void k_gather_fill(long n, const int* __restrict__ sel, double* __restrict__ dst, double v)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[sel[q]] = v;
}

// values computed on the host (SHX_MMS_HOST) into their places: point sel[q]
// of D takes v[q]; cut point i takes v[i] where it is Dirichlet (set) or adds
// it where it is not (add)
__global__ void k_scatter(long n, const int* __restrict__ sel, const double* __restrict__ v, double* __restrict__ dst)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[sel[q]] = v[q];
}

__global__ void k_r_set_dirichlet(long n, const signed char* __restrict__ dirich, const double* __restrict__ v,
                                  double* __restrict__ R)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && dirich[i]) R[i] = v[i];
}

__global__ void k_r_add_free(long n, const signed char* __restrict__ dirich, const double* __restrict__ v,
                             double* __restrict__ R)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !dirich[i]) R[i] += v[i];
}

// heat.cpp:140-146: the source at the cut points of non-Dirichlet bodies
__global__ void k_src_R_add(DevSys s, long n, const double* __restrict__ pos,
                            const signed char* __restrict__ dirich, double* __restrict__ R, double t)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || dirich[i]) return;
    R[i] += heat_source(s, t, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

// heat.cpp:154-158: the cut points of Dirichlet bodies take zero
__global__ void k_zero_dirichlet(long n, const signed char* __restrict__ dirich, double* __restrict__ R)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && dirich[i]) R[i] = 0.0;
}

// heat.cpp:318-346: the gradient component dir of the solution on one
// Neumann face
__global__ void k_neumann_data(DevSys s, long n, const int* __restrict__ sel, double* __restrict__ neu,
                               int dir, double t)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long idx = sel[q];
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    neu[idx] = mms_grad(s.mms_kind, s.mms_dims, s.mms_n, s.mms, t, s.x[i], s.y[j], s.z[k], dir);
}

// slot_ops.hpp:23-59
__global__ void k_assign_lc(long n, double* __restrict__ d, const double* __restrict__ s0, double c,
                            const double* __restrict__ r)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s0[i] + c * r[i];
}

__global__ void k_accumulate(long n, double* __restrict__ d, double c, const double* __restrict__ r)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] += c * r[i];
}

__global__ void k_copy(long n, double* __restrict__ d, const double* __restrict__ s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

// compute_scalar_stats partials over the fluid gather: per workgroup
// (max |u - exact|, its index, min u, max u)
__global__ __launch_bounds__(256) void k_stats(DevSys s, long n, const int* __restrict__ fluid,
                                               const double* __restrict__ u, const double* __restrict__ exF,
                                               double t,
                                               double* __restrict__ part_err,
                                               long* __restrict__ part_idx,
                                               double* __restrict__ part_min,
                                               double* __restrict__ part_max)
{
    __shared__ double se[256], smin[256], smax[256];
    __shared__ long si[256];
    double e = -DBL_MAX, mn = DBL_MAX, mx = -DBL_MAX;
    long ei = 0x7fffffffffffffffL;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long idx = fluid[q];
        int i, j, k;
        point_coords(idx, s.n1, s.n2, i, j, k);
        const double v = u[idx];
        const double err = fabs(v - (exF ? exF[q] : exact_at(s, s.x[i], s.y[j], s.z[k], t)));
        // strict comparisons, as the reference's MaxLoc: a NaN error never
        // wins (scalar_system_utils.hpp:102-113); ties keep the first point
        if (err > e || (err == e && idx < ei)) {
            e = err;
            ei = idx;
        }
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    se[threadIdx.x] = e;
    si[threadIdx.x] = ei;
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const int o = threadIdx.x + w;
            const double a = se[o], b = se[threadIdx.x];
            if (a > b || (a == b && si[o] < si[threadIdx.x])) {
                se[threadIdx.x] = se[o];
                si[threadIdx.x] = si[o];
            }
            smin[threadIdx.x] = smin[o] < smin[threadIdx.x] ? smin[o] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[o] > smax[threadIdx.x] ? smax[o] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_err[blockIdx.x] = se[0];
        part_idx[blockIdx.x] = si[0];
        part_min[blockIdx.x] = smin[0];
        part_max[blockIdx.x] = smax[0];
    }
}

inline unsigned grid_for(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---------------------------------------------------------------- system state

namespace {

int check(shx_ctx* ctx, const char* what)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, what);
}

int alloc_d(shx_ctx* ctx, double** p, long n)
{
    hipError_t e = hipMalloc((void**)p, (size_t)(n > 0 ? n : 1) * sizeof(double));
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "hipMalloc");
}

RPtrs rptrs(const shx_field* f)
{
    RPtrs r;
    for (int c = 0; c < 3; ++c) r.r[c] = f->R[c];
    return r;
}

// ---- SHX_MMS_HOST: the manufactured solution's functions called on the host
// (lua_mms.hpp:12-79), their values copied to the device
enum { kHmCall = 0, kHmSource = 1, kHmGrad = 2 };  // kHmGrad + d: component d of the gradient

double hm_value(const shx_system* s, int what, double t, const double loc[3])
{
    const shx_mms_fns& f = s->hm->f;
    if (what == kHmCall) return f.call(f.ctx, t, loc);
    if (what == kHmSource) return f.ddt(f.ctx, t, loc) - s->dev.k * f.lap(f.ctx, t, loc);  // heat.cpp:111-117
    double g[3] = {0.0, 0.0, 0.0};
    f.grad(f.ctx, t, loc, g);
    return g[what - kHmGrad];
}

// values at the mesh points sel[0..n) (all N points in index order when sel is
// null) into the host staging
void hm_mesh(shx_system* s, int what, double t, const int* sel, long n)
{
    HostMms& h = *s->hm;
    h.h.resize((size_t)n);
    const long plane = (long)s->n[1] * s->n[2];
    for (long q = 0; q < n; ++q) {
        const long idx = sel ? sel[q] : q;
        const int i = (int)(idx / plane), rem = (int)(idx - (long)i * plane);
        const double loc[3] = {h.xyz[0][i], h.xyz[1][rem / s->n[2]], h.xyz[2][rem % s->n[2]]};
        h.h[(size_t)q] = hm_value(s, what, t, loc);
    }
}

// values at the cut points of R(c) into the host staging
void hm_cut(shx_system* s, int c, int what, double t)
{
    HostMms& h = *s->hm;
    h.h.resize((size_t)s->nR[c]);
    for (long i = 0; i < s->nR[c]; ++i) h.h[(size_t)i] = hm_value(s, what, t, &h.pos[c][(size_t)(3 * i)]);
}

// the host staging to dst (device), completed before returning
int hm_put(shx_system* s, double* dst)
{
    HostMms& h = *s->hm;
    if (h.h.empty()) return SHX_OK;
    hipError_t e = hipMemcpyAsync(dst, h.h.data(), h.h.size() * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    return e == hipSuccess ? SHX_OK : hip_fail(s->ctx, e, "host solution values h2d");
}

int r_exact_all(shx_system* s, double* const R[3], double t)
{
    if (s->hm) {
        for (int c = 0; c < 3; ++c) {
            if (s->nR[c] == 0) continue;
            hm_cut(s, c, kHmCall, t);
            int rc = hm_put(s, s->hm->d);
            if (rc != SHX_OK) return rc;
            k_r_set_dirichlet<<<grid_for(s->nR[c]), 256, 0, s->ctx->stream>>>(s->nR[c], s->rdir[c], s->hm->d, R[c]);
            rc = check(s->ctx, "cut-point boundary kernel");
            if (rc == SHX_OK) rc = sync(s->ctx);
            if (rc != SHX_OK) return rc;
        }
        return SHX_OK;
    }
    for (int c = 0; c < 3; ++c) {
        if (s->nR[c] == 0) continue;
        k_r_exact<<<grid_for(s->nR[c]), 256, 0, s->ctx->stream>>>(s->dev, s->nR[c], s->rpos[c],
                                                                  s->rdir[c], R[c], t);
    }
    return check(s->ctx, "cut-point boundary kernel");
}

int face_exact(shx_system* s, double* D, double t)
{
    if (s->nfaces == 0) return SHX_OK;
    if (s->hm) {
        hm_mesh(s, kHmCall, t, s->hm->faces.data(), s->nfaces);
        int rc = hm_put(s, s->hm->d);
        if (rc != SHX_OK) return rc;
        k_scatter<<<grid_for(s->nfaces), 256, 0, s->ctx->stream>>>(s->nfaces, s->faces, s->hm->d, D);
        rc = check(s->ctx, "face boundary kernel");
        return rc == SHX_OK ? sync(s->ctx) : rc;
    }
    k_face_exact<<<grid_for(s->nfaces), 256, 0, s->ctx->stream>>>(s->dev, s->nfaces, s->faces, D, t);
    return check(s->ctx, "face boundary kernel");
}

// Fused-form registers: the accumulator, the stage ping-pong (FP64, or FP32
// for the mixed form) and the stage's cut-point values (FP64).
int ensure_fused(shx_system* s)
{
    int rc = SHX_OK;
    if (!s->accD) {
        rc = alloc_d(s->ctx, &s->accD, s->N);
        for (int b = 0; b < 2 && rc == SHX_OK; ++b)
            for (int c = 0; c < 3 && rc == SHX_OK; ++c) rc = alloc_d(s->ctx, &s->stageR[b][c], s->nR[c]);
        if (rc == SHX_OK && s->nfree > 0) rc = alloc_d(s->ctx, &s->accF, s->nfree);
    }
    for (int b = 0; b < 2 && rc == SHX_OK; ++b) {
        if (s->precision == SHX_F64 && !s->stageD[b]) rc = alloc_d(s->ctx, &s->stageD[b], s->N);
        if (s->precision == SHX_F32_COMP && !s->stageF[b]) {
            const hipError_t e = hipMalloc((void**)&s->stageF[b], sizeof(float) * (size_t)(s->N > 0 ? s->N : 1));
            if (e != hipSuccess) rc = hip_fail(s->ctx, e, "hipMalloc");
        }
    }
    return rc;
}

// Full-length B (rows = N) rebuilt from the segment table in the builder's
// (row, col, value) order, src/matrices/csr.cpp:10-35.
int build_bfull(shx_system* s, int c)
{
    shx_derivative* d = s->d[c];
    struct E {
        long row, col;
        double w;
    };
    std::vector<E> e;
    const long stride = d->dev.stride;
    const int n1 = s->n[1], n2 = s->n[2];
    const int nfast = c == 2 ? n1 : n2;
    for (const shx_line_block& b : d->h_blocks) {
        const int slow = b.pair / nfast, fast = b.pair % nfast;
        long base;
        if (c == 0) base = (long)slow * n2 + fast;                     // (j, k)
        else if (c == 1) base = (long)slow * n1 * n2 + fast;          // (i, k)
        else base = ((long)slow * n1 + fast) * n2;                    // (i, j)
        const int total = b.nl + b.ni + b.nr;
        if (b.lobj >= 0)
            for (int r = 0; r < b.nl; ++r)
                e.push_back({base + (long)(b.lo + r) * stride, b.lobj, d->h_pool[b.lbw + r]});
        if (b.robj >= 0)
            for (int r = 0; r < b.nr; ++r)
                e.push_back({base + (long)(b.lo + total - b.nr + r) * stride, b.robj,
                             d->h_pool[b.rbw + r]});
    }
    std::sort(e.begin(), e.end(), [](const E& a, const E& b) {
        if (a.row != b.row) return a.row < b.row;
        if (a.col != b.col) return a.col < b.col;
        return a.w < b.w;
    });
    std::vector<int64_t> u(s->N + 1, 0), v(e.size());
    std::vector<double> w(e.size());
    for (size_t q = 0; q < e.size(); ++q) {
        u[e[q].row + 1]++;
        v[q] = e[q].col;
        w[q] = e[q].w;
    }
    for (long r = 0; r < s->N; ++r) u[r + 1] += u[r];
    return shx_csr_create(s->ctx, s->N, u.data(), v.data(), w.data(), &s->Bfull[c]);
}

int ensure_unfused(shx_system* s)
{
    if (s->unfused_ready) return SHX_OK;
    int rc = SHX_OK;
    if (s->kind != SHX_SYS_HEAT) {
        for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
            rc = alloc_d(s->ctx, &s->gG[c][0], s->N);
            for (int r = 0; r < 3 && rc == SHX_OK; ++r) rc = alloc_d(s->ctx, &s->gG[c][r + 1], s->nR[r]);
            rc = rc == SHX_OK ? alloc_d(s->ctx, &s->du[c][0], s->N) : rc;
            for (int r = 0; r < 3 && rc == SHX_OK; ++r) rc = alloc_d(s->ctx, &s->du[c][r + 1], s->nR[r]);
        }
        if (rc != SHX_OK) return rc;
        k_gG_D<<<grid_for(s->N), 256, 0, s->ctx->stream>>>(s->dev, s->N, s->gG[0][0], s->gG[1][0],
                                                          s->gG[2][0]);
        // gG on cut points: -g at those of non-Dirichlet bodies, zero at
        // Dirichlet ones
        for (int r = 0; r < 3; ++r)
            if (s->nR[r] > 0)
                k_gG_R<<<grid_for(s->nR[r]), 256, 0, s->ctx->stream>>>(s->dev, s->nR[r], s->rpos[r], s->rdir[r],
                                                                      s->gG[0][r + 1], s->gG[1][r + 1],
                                                                      s->gG[2][r + 1]);
        rc = rc == SHX_OK ? check(s->ctx, "gG kernel") : rc;
    } else {
        rc = alloc_d(s->ctx, &s->srcD, s->N);
    }
    for (int c = 0; c < 3 && rc == SHX_OK; ++c)
        if (s->d[c] && s->n[c] > 1) rc = build_bfull(s, c);
    if (rc == SHX_OK) rc = sync(s->ctx);
    s->unfused_ready = rc == SHX_OK;
    return rc;
}

int rhs_unfused_launch(shx_system* s, const shx_field* u, shx_field* rhs, double t);

int rhs_fused_launch(shx_system* s, const shx_field* u, shx_field* rhs, double t)
{
    // the point-wise form covers Dirichlet bodies on grids without Neumann
    // faces; other systems take the reference's sequence
    if (s->general || s->hm) return rhs_unfused_launch(s, u, rhs, t);
    const Tiling& tl = s->tiling;
    hipStream_t st = s->ctx->stream;
    switch (s->family * 10 + s->width) {
#define SHX_RHS_CASE(K, W)                                                                      \
    case K * 10 + W:                                                                            \
        k_rhs<K, W><<<tl.nblocks, kBlock, 0, st>>>(tl, s->dev, s->ops, u->D, rptrs(u), rhs->D, t); \
        break;
        SHX_RHS_CASE(SHX_SYS_HEAT, 3) SHX_RHS_CASE(SHX_SYS_HEAT, 5)
        SHX_RHS_CASE(kHeatGauss, 3) SHX_RHS_CASE(kHeatGauss, 5)
        SHX_RHS_CASE(SHX_SYS_WAVE_LINEAR, 3) SHX_RHS_CASE(SHX_SYS_WAVE_LINEAR, 5)
        SHX_RHS_CASE(SHX_SYS_WAVE_RADIAL, 3) SHX_RHS_CASE(SHX_SYS_WAVE_RADIAL, 5)
#undef SHX_RHS_CASE
    default:
        return fail(s->ctx, SHX_EUNSUPPORTED, "rhs: interior width must be 3 or 5");
    }
    int rc = check(s->ctx, "rhs kernel");
    // every body is Dirichlet: the cut-point rows of the rhs are zero
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) rc = launch_fill(s->ctx, rhs->R[c], 0.0, s->nR[c]);
    return rc;
}

int rhs_unfused_launch(shx_system* s, const shx_field* u, shx_field* rhs, double t)
{
    int rc = ensure_unfused(s);
    if (rc != SHX_OK) return rc;
    shx_ctx* ctx = s->ctx;
    hipStream_t st = ctx->stream;
    if (s->kind != SHX_SYS_HEAT) {
        // gradient: zero 12 buffers, then O and B per direction (gradient.hpp:35-122)
        for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
            rc = launch_fill(ctx, s->du[c][0], 0.0, s->N);
            for (int r = 0; r < 3 && rc == SHX_OK; ++r) rc = launch_fill(ctx, s->du[c][r + 1], 0.0, s->nR[r]);
        }
        for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
            if (!s->d[c] || s->n[c] < 2) continue;
            // derivative.cpp:479-485: the cut-point rows of non-Dirichlet bodies
            shx_field duc{};
            duc.D = s->du[c][0];
            for (int r = 0; r < 3; ++r) duc.R[r] = s->du[c][r + 1];
            rc = launch_cut_rows(s->d[c], u, &duc);
            if (rc == SHX_OK) rc = launch_derivative(s->d[c], u->D, nullptr, s->du[c][0], false, false);
            if (rc == SHX_OK) rc = launch_csr_full(s->Bfull[c], u->R[c], s->du[c][0]);
        }
        if (rc != SHX_OK) return rc;
        k_dot3<<<grid_for(s->N), 256, 0, st>>>(s->N, s->gG[0][0], s->du[0][0], s->gG[1][0], s->du[1][0],
                                               s->gG[2][0], s->du[2][0], rhs->D);
        for (int r = 0; r < 3; ++r)
            if (s->nR[r] > 0)
                k_dot3<<<grid_for(s->nR[r]), 256, 0, st>>>(s->nR[r], s->gG[0][r + 1], s->du[0][r + 1],
                                                           s->gG[1][r + 1], s->du[1][r + 1],
                                                           s->gG[2][r + 1], s->du[2][r + 1], rhs->R[r]);
        return check(ctx, "dot kernel");
    }
    // heat: laplacian (zero, O+B per direction, plus_eq), scale, source, fills
    rc = launch_fill(ctx, rhs->D, 0.0, s->N);
    for (int r = 0; r < 3 && rc == SHX_OK; ++r) rc = launch_fill(ctx, rhs->R[r], 0.0, s->nR[r]);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
        if (!s->d[c] || s->n[c] < 2) continue;
        // laplacian.cpp:48-57 with derivative.cpp:475-519: cut-point rows,
        // O and B, then N nu with the Neumann face data
        rc = launch_cut_rows(s->d[c], u, rhs);
        if (rc == SHX_OK) rc = launch_derivative(s->d[c], u->D, nullptr, rhs->D, true, false);
        if (rc == SHX_OK) rc = launch_csr_full(s->Bfull[c], u->R[c], rhs->D);
        if (rc == SHX_OK && s->neuD) rc = launch_neumann(s->d[c], s->neuD, rhs->D);
    }
    if (rc != SHX_OK) return rc;
    k_scale<<<grid_for(s->N), 256, 0, st>>>(s->N, rhs->D, s->dev.k);
    for (int r = 0; r < 3; ++r)
        if (s->nR[r] > 0) k_scale<<<grid_for(s->nR[r]), 256, 0, st>>>(s->nR[r], rhs->R[r], s->dev.k);
    if (s->hm) {
        hm_mesh(s, kHmSource, t, nullptr, s->N);
        rc = hm_put(s, s->srcD);
        if (rc != SHX_OK) return rc;
    } else {
        k_src_D<<<grid_for(s->N), 256, 0, st>>>(s->dev, s->N, s->srcD, t);
    }
    if (s->nfluid) k_gather_add<<<grid_for(s->nfluid), 256, 0, st>>>(s->nfluid, s->fluid, s->srcD, rhs->D);
    for (int r = 0; r < 3; ++r) {
        if (s->nR[r] == 0 || !s->any_free_r) continue;
        if (s->hm) {
            hm_cut(s, r, kHmSource, t);
            rc = hm_put(s, s->hm->d);
            if (rc != SHX_OK) return rc;
            k_r_add_free<<<grid_for(s->nR[r]), 256, 0, st>>>(s->nR[r], s->rdir[r], s->hm->d, rhs->R[r]);
            rc = check(ctx, "heat source kernels");
            if (rc == SHX_OK) rc = sync(ctx);
            if (rc != SHX_OK) return rc;
        } else {
            k_src_R_add<<<grid_for(s->nR[r]), 256, 0, st>>>(s->dev, s->nR[r], s->rpos[r], s->rdir[r], rhs->R[r], t);
        }
    }
    if (s->nfaces) k_gather_fill<<<grid_for(s->nfaces), 256, 0, st>>>(s->nfaces, s->faces, rhs->D, 0.0);
    for (int r = 0; r < 3; ++r)
        if (s->nR[r] > 0) k_zero_dirichlet<<<grid_for(s->nR[r]), 256, 0, st>>>(s->nR[r], s->rdir[r], rhs->R[r]);
    return check(ctx, "heat source kernels");
}

int update_boundary_launch(shx_system* s, shx_field* u, double t)
{
    int rc = face_exact(s, u->D, t);
    if (rc == SHX_OK) rc = r_exact_all(s, u->R, t);
    // heat.cpp:318-346: Neumann face data, face by face in direction order
    // (an edge point shared by two faces keeps the later direction's value)
    if (rc == SHX_OK && s->neuD) {
        for (int f = 0; f < 6; ++f) {
            const long n = s->neu_off[f + 1] - s->neu_off[f];
            if (n > 0 && s->hm) {
                hm_mesh(s, kHmGrad + f / 2, t, s->hm->neu.data() + s->neu_off[f], n);
                rc = hm_put(s, s->hm->d);
                if (rc != SHX_OK) return rc;
                k_scatter<<<grid_for(n), 256, 0, s->ctx->stream>>>(n, s->neu_faces + s->neu_off[f], s->hm->d, s->neuD);
                rc = check(s->ctx, "neumann data kernel");
                if (rc == SHX_OK) rc = sync(s->ctx);
                if (rc != SHX_OK) return rc;
            } else if (n > 0) {
                k_neumann_data<<<grid_for(n), 256, 0, s->ctx->stream>>>(s->dev, n, s->neu_faces + s->neu_off[f],
                                                                       s->neuD, f / 2, t);
            }
        }
        rc = check(s->ctx, "neumann data kernel");
    }
    return rc;
}

int slot_assign_lc(shx_system* s, shx_field* d, const shx_field* s0, double c, const shx_field* r)
{
    hipStream_t st = s->ctx->stream;
    k_assign_lc<<<grid_for(s->N), 256, 0, st>>>(s->N, d->D, s0->D, c, r->D);
    for (int q = 0; q < 3; ++q)
        if (s->nR[q] > 0) k_assign_lc<<<grid_for(s->nR[q]), 256, 0, st>>>(s->nR[q], d->R[q], s0->R[q], c, r->R[q]);
    return check(s->ctx, "assign_lc kernel");
}

int slot_accumulate(shx_system* s, shx_field* d, double c, const shx_field* r)
{
    hipStream_t st = s->ctx->stream;
    k_accumulate<<<grid_for(s->N), 256, 0, st>>>(s->N, d->D, c, r->D);
    for (int q = 0; q < 3; ++q)
        if (s->nR[q] > 0) k_accumulate<<<grid_for(s->nR[q]), 256, 0, st>>>(s->nR[q], d->R[q], c, r->R[q]);
    return check(s->ctx, "accumulate kernel");
}

int slot_copy(shx_system* s, shx_field* d, const shx_field* src)
{
    hipStream_t st = s->ctx->stream;
    k_copy<<<grid_for(s->N), 256, 0, st>>>(s->N, d->D, src->D);
    for (int q = 0; q < 3; ++q)
        if (s->nR[q] > 0) k_copy<<<grid_for(s->nR[q]), 256, 0, st>>>(s->nR[q], d->R[q], src->R[q]);
    return check(s->ctx, "copy kernel");
}

int zero_slot(shx_system* s, shx_field* f) { return zero_field(s->ctx, f, s->n, s->nR); }

constexpr double kRki[4] = {0.0, 0.5, 0.5, 1.0};
constexpr double kRkf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

// rk4.cpp:16-47, one kernel per reference operation
int rk4_unfused(shx_system* s, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs,
                double t, double dt)
{
    int rc = zero_slot(s, rk);
    if (rc == SHX_OK) rc = zero_slot(s, srhs);
    if (rc == SHX_OK) rc = slot_copy(s, out, u0);
    for (int i = 0; i < 4 && rc == SHX_OK; ++i) {
        if (i > 0) {
            rc = slot_assign_lc(s, out, u0, dt * kRki[i], srhs);
            if (rc == SHX_OK) rc = update_boundary_launch(s, out, t + dt * kRki[i]);
        }
        if (rc == SHX_OK) rc = rhs_unfused_launch(s, out, srhs, t + dt * kRki[i]);
        if (rc == SHX_OK) rc = slot_accumulate(s, rk, dt * kRkf[i], srhs);
    }
    if (rc == SHX_OK) rc = slot_assign_lc(s, out, u0, 1.0, rk);
    if (rc == SHX_OK) rc = update_boundary_launch(s, out, t + dt);
    return rc;
}

template <int KIND>
int launch_stage(shx_system* s, int stage, const double* usD, RPtrs usR, const double* u0,
                 double* acc, double* out, double ts, double c_acc, double c_next, double tn)
{
    const Tiling& tl = s->tiling;
    hipStream_t st = s->ctx->stream;
    const int key = stage * 10 + s->width;
    switch (key) {
#define SHX_STAGE_CASE(S, W)                                                                  \
    case S * 10 + W:                                                                          \
        k_rk4_stage<KIND, S, W><<<tl.nblocks, kBlock, 0, st>>>(tl, s->dev, s->ops, usD, usR, \
                                                             u0, acc, out, ts, c_acc, c_next, tn); \
        break;
        SHX_STAGE_CASE(0, 3) SHX_STAGE_CASE(1, 3) SHX_STAGE_CASE(2, 3) SHX_STAGE_CASE(3, 3)
        SHX_STAGE_CASE(0, 5) SHX_STAGE_CASE(1, 5) SHX_STAGE_CASE(2, 5) SHX_STAGE_CASE(3, 5)
#undef SHX_STAGE_CASE
    default:
        return fail(s->ctx, SHX_EUNSUPPORTED, "rk4 stage: interior width must be 3 or 5");
    }
    return check(s->ctx, "rk4 stage kernel");
}

// The x-marching form lists special points with 32-bit indices.
bool march_ok(const shx_system* s)
{
    return (s->width == 3 || s->width == 5) && s->N < (1L << 32);
}

template <typename T>
int upload_vec(shx_ctx* ctx, T** dst, const std::vector<T>& v)
{
    *dst = nullptr;
    if (v.empty()) return SHX_OK;
    hipError_t e = hipMalloc((void**)dst, v.size() * sizeof(T));
    if (e == hipSuccess) e = hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, "special-point upload");
}

// Class byte of every point and the special-point records (built once).
int ensure_classes(shx_system* s)
{
    if (s->classes_ready) return SHX_OK;
    if (!march_ok(s)) return fail(s->ctx, SHX_EUNSUPPORTED, "x-marching form: unsupported system");
    hipError_t e = hipMalloc((void**)&s->cls, (size_t)s->N);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "classes alloc");
    launch_classify(s->ctx->stream, s->dev, s->ops, s->cls, s->N);
    int rc = check(s->ctx, "classify kernel");
    std::vector<unsigned char> h((size_t)s->N);
    if (rc == SHX_OK) {
        e = hipMemcpyAsync(h.data(), s->cls, (size_t)s->N, hipMemcpyDeviceToHost, s->ctx->stream);
        if (e != hipSuccess) rc = hip_fail(s->ctx, e, "classes d2h");
    }
    if (rc == SHX_OK) rc = sync(s->ctx);
    if (rc != SHX_OK) return rc;
    // Raw special points become face points, closure-pattern points of the
    // x-marching tiles, or special points with a record.  Rows other than
    // interior ones go to per-direction tables, deduplicated (grid-wall
    // closures repeat from line to line).
    const int n1 = s->n[1], n2 = s->n[2];
    const int hw = s->width / 2;
    // 0 = no row, 1 = interior row, 2 = other row (e filled); the segment
    // holding row q is found as the device slow path finds it
    auto find_row = [&](int c, int pair, int q, SpecialRow& e) -> int {
        const shx_derivative* d = s->d[c];
        for (int b = d->h_pair_ptr[pair]; b < d->h_pair_ptr[pair + 1]; ++b) {
            const shx_line_block& blk = d->h_blocks[b];
            const int lr = q - blk.lo;
            if (lr < 0) break;
            const int total = blk.nl + blk.ni + blk.nr;
            if (lr >= total) continue;
            if (lr >= blk.nl && lr < blk.nl + blk.ni) return 1;
            e = SpecialRow{};
            if (lr < blk.nl) {
                e.off = blk.lcol0 - q;
                e.ncols = blk.lcols;
                e.coef = blk.lcoef + lr * blk.lcols;
            } else {
                const int rr = lr - blk.nl - blk.ni;
                e.off = blk.rcol0 - q;
                e.ncols = blk.rcols;
                e.coef = blk.rcoef + rr * blk.rcols;
            }
            e.bobj0 = e.bw0 = e.bobj1 = e.bw1 = -1;
            if (blk.lobj >= 0 && lr < blk.nl) {
                e.bobj0 = blk.lobj;
                e.bw0 = blk.lbw + lr;
                e.nb = 1;
            }
            if (blk.robj >= 0 && lr >= total - blk.nr) {
                if (e.nb == 0) {
                    e.bobj0 = blk.robj;
                    e.bw0 = blk.rbw + (lr - (total - blk.nr));
                } else {
                    e.bobj1 = blk.robj;
                    e.bw1 = blk.rbw + (lr - (total - blk.nr));
                }
                e.nb++;
            }
            return 2;
        }
        return 0;
    };
    auto rows_of = [&](long idx, int kind[3], SpecialRow e[3]) {
        const int i = (int)(idx / ((long)n1 * n2));
        const int rem = (int)(idx - (long)i * n1 * n2);
        const int j = rem / n2, k = rem - (rem / n2) * n2;
        for (int c = 0; c < 3; ++c) {
            kind[c] = 1;
            if (!s->d[c] || !s->ops.d[c].present) continue;
            const int pair = c == 0 ? j * n2 + k : c == 1 ? i * n2 + k : i * n1 + j;
            const int q = c == 0 ? i : c == 1 ? j : k;
            kind[c] = find_row(c, pair, q, e[c]);
        }
    };
    // a closure row without B terms whose columns [q+off, q+off+n) lie in the
    // mesh and in what the tile holds: the register queue for x, the LDS
    // plane of every tile height for y (the 4-row tile is the smallest) and z
    auto reachable = [&](int c, const SpecialRow& e, int q) {
        if (e.nb != 0 || e.ncols < 1 || e.ncols > kMaxClosureCols) return false;
        const int first = q + e.off, last = q + e.off + e.ncols - 1;
        // x rows: from the register queue when they fit, else the tile reads
        // the stage input at the row's planes
        if (c == 0) return first >= 0 && last < s->n[0];
        const int tile = c == 1 ? 4 : 64, n = c == 1 ? n1 : n2;
        const int lo = std::max(0, (q / tile) * tile - hw);
        const int hi = std::min(n, (q / tile) * tile + tile + hw);
        return first >= lo && last < hi;
    };
    struct Cand {
        long idx;
        int rid[3];
    };
    std::vector<bool> raw_face((size_t)s->N, false);
    std::vector<Cand> cands;
    std::vector<ClosureRow> allrows[3];
    std::vector<long> freq[3];
    std::map<std::array<int, 3>, int> rowid[3];
    for (long idx = 0; idx < s->N; ++idx) {
        const unsigned char raw = h[(size_t)idx];
        if (!(raw & kRawSpecial)) continue;
        raw_face[(size_t)idx] = (raw & kRawFace) != 0;
        h[(size_t)idx] = kSpecial;
        if ((raw & kRawFace) || !(raw & kRawFluid)) continue;
        int kind[3];
        SpecialRow e[3];
        rows_of(idx, kind, e);
        const int i = (int)(idx / ((long)n1 * n2));
        const int rem = (int)(idx - (long)i * n1 * n2);
        const int qs[3] = {i, rem / n2, rem % n2};
        bool ok = true;
        for (int c = 0; c < 3 && ok; ++c)
            ok = kind[c] == 1 || (kind[c] == 2 && reachable(c, e[c], qs[c]));
        if (!ok) continue;
        Cand cd{idx, {-1, -1, -1}};
        for (int c = 0; c < 3; ++c) {
            if (kind[c] != 2) continue;
            const std::array<int, 3> key{e[c].off, e[c].ncols, e[c].coef};
            auto it = rowid[c].find(key);
            if (it == rowid[c].end()) {
                it = rowid[c].emplace(key, (int)allrows[c].size()).first;
                ClosureRow r{};
                r.off = e[c].off;
                r.n = e[c].ncols;
                for (int m = 0; m < r.n; ++m) r.w[m] = s->d[c]->h_pool[(size_t)e[c].coef + m];
                allrows[c].push_back(r);
                freq[c].push_back(0);
            }
            cd.rid[c] = it->second;
            freq[c][it->second]++;
        }
        cands.push_back(cd);
    }
    // rank the rows of each direction by use; pick table sizes (nx, ny, nz),
    // (nx+1)(ny+1)(nz+1) <= kMaxPatterns, covering the most points
    constexpr int R1 = kMaxClosureRows + 1;  // rank R1 = not in any table
    std::vector<int> rank[3];
    for (int c = 0; c < 3; ++c) {
        std::vector<int> order(allrows[c].size());
        for (size_t r = 0; r < order.size(); ++r) order[r] = (int)r;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return freq[c][x] > freq[c][y]; });
        rank[c].assign(order.size(), R1);
        for (size_t r = 0; r < order.size() && (int)r < kMaxClosureRows; ++r) rank[c][order[r]] = (int)r + 1;
    }
    auto rk = [&](const Cand& cd, int c) { return cd.rid[c] < 0 ? 0 : rank[c][cd.rid[c]]; };
    std::vector<long> hist((size_t)(R1 + 1) * (R1 + 1) * (R1 + 1), 0);
    auto at = [&](int x, int y, int z) -> long& { return hist[((size_t)x * (R1 + 1) + y) * (R1 + 1) + z]; };
    for (const Cand& cd : cands) at(rk(cd, 0), rk(cd, 1), rk(cd, 2))++;
    for (int x = 0; x <= R1; ++x)  // inclusive prefix sums
        for (int y = 0; y <= R1; ++y)
            for (int z = 0; z <= R1; ++z) {
                long v = at(x, y, z);
                if (x) v += at(x - 1, y, z);
                if (y) v += at(x, y - 1, z);
                if (z) v += at(x, y, z - 1);
                if (x && y) v -= at(x - 1, y - 1, z);
                if (x && z) v -= at(x - 1, y, z - 1);
                if (y && z) v -= at(x, y - 1, z - 1);
                if (x && y && z) v += at(x - 1, y - 1, z - 1);
                at(x, y, z) = v;
            }
    int nr[3] = {0, 0, 0};
    long best = -1;
    const int lim[3] = {std::min<int>(kMaxClosureRows, (int)allrows[0].size()),
                        std::min<int>(kMaxClosureRows, (int)allrows[1].size()),
                        std::min<int>(kMaxClosureRows, (int)allrows[2].size())};
    for (int x = 0; x <= lim[0]; ++x)
        for (int y = 0; y <= lim[1]; ++y)
            for (int z = 0; z <= lim[2]; ++z) {
                if ((x + 1) * (y + 1) * (z + 1) > kMaxPatterns) continue;
                const long cov = at(x, y, z);
                if (cov > best) {
                    best = cov;
                    nr[0] = x;
                    nr[1] = y;
                    nr[2] = z;
                }
            }
    for (const Cand& cd : cands) {
        const int rx = rk(cd, 0), ry = rk(cd, 1), rz = rk(cd, 2);
        if (rx > nr[0] || ry > nr[1] || rz > nr[2]) continue;
        h[(size_t)cd.idx] = (unsigned char)(kPattern0 + (rx * (nr[1] + 1) + ry) * (nr[2] + 1) + rz);
    }
    std::vector<ClosureRow> table((size_t)3 * kMaxClosureRows, ClosureRow{});
    for (int c = 0; c < 3; ++c) {
        s->nrows[c] = nr[c];
        s->fixw[c] = 1;
        for (size_t r = 0; r < allrows[c].size(); ++r)
            if (rank[c][r] <= nr[c]) {
                table[(size_t)c * kMaxClosureRows + rank[c][r] - 1] = allrows[c][r];
                if (allrows[c][r].n != s->width) s->fixw[c] = 0;
            }
    }
    // records of the remaining special points, in index order.  A Dirichlet
    // face point needs no rows (its right-hand side is zero: gG = 0 or the
    // heat face rule) and takes exact(t_next): the faces are special items
    // of their own, enumerated plane by plane
    std::vector<SpecialRec> recs;
    s->nface_points = 0;
    s->nface_items = 0;
    for (int f = 0; f < 6; ++f)
        if (s->dev.bc[f] == SHX_BC_DIRICHLET)
            s->nface_items += f < 2 ? (long)n1 * n2 : f < 4 ? (long)s->n[0] * n2 : (long)s->n[0] * n1;
    std::vector<SpecialRow> rows[3];
    std::map<std::array<int, 8>, unsigned> seen[3];
    for (long idx = 0; idx < s->N; ++idx) {
        if (h[(size_t)idx] != kSpecial) continue;
        SpecialRec r;
        r.idx = (unsigned)idx;
        if (raw_face[(size_t)idx]) {  // a face item, not a record
            ++s->nface_points;
            continue;
        }
        int kind[3];
        SpecialRow e[3];
        rows_of(idx, kind, e);
        for (int c = 0; c < 3; ++c) {
            r.ref[c] = kind[c] == 1 ? kRowInterior : kRowNone;
            if (kind[c] != 2) continue;
            const SpecialRow& x = e[c];
            const std::array<int, 8> key{x.off, x.ncols, x.coef, x.nb, x.bobj0, x.bw0, x.bobj1, x.bw1};
            auto it = seen[c].find(key);
            if (it == seen[c].end()) {
                it = seen[c].emplace(key, (unsigned)rows[c].size()).first;
                rows[c].push_back(x);
            }
            r.ref[c] = it->second;
        }
        if (!s->d[0] || !s->ops.d[0].present) r.ref[0] = kRowNone;
        if (!s->d[1] || !s->ops.d[1].present) r.ref[1] = kRowNone;
        if (!s->d[2] || !s->ops.d[2].present) r.ref[2] = kRowNone;
        recs.push_back(r);
    }
    s->nspecial = (long)recs.size();
    e = hipMemcpy(s->cls, h.data(), (size_t)s->N, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "classes h2d");
    int rc2 = upload_vec(s->ctx, &s->special, recs);
    if (rc2 == SHX_OK) rc2 = upload_vec(s->ctx, &s->crows, table);
    for (int c = 0; c < 3 && rc2 == SHX_OK; ++c) rc2 = upload_vec(s->ctx, &s->srows[c], rows[c]);
    if (rc2 != SHX_OK) return rc2;
    s->classes_ready = true;
    return SHX_OK;
}

// Tile and special-point arguments of one stage of the x-marching form.
int march_args(shx_system* s, const void* usD, RPtrs usR, const double* u0, RPtrs u0R, double* acc,
               double* accF, void* out, double* const outR[3], MarchArgs& a, SpecialArgs& sp, long& nitems)
{
    a.s = s->dev;
    a.ops = s->ops;
    a.us = usD;
    a.u0 = u0;
    a.acc = acc;
    a.out = out;
    a.cls = s->cls;
    a.rows = s->crows;
    for (int c = 0; c < 3; ++c) {
        a.nrows[c] = s->nrows[c];
        a.fixw[c] = s->fixw[c];
    }
    a.xchunk = s->march_xchunk;
    // two points per lane along z when every pair is aligned (even n2)
    a.vec = s->march_vec != 0 ? s->march_vec : (s->n[2] % 2 == 0 ? 2 : 1);
    a.nbz = (s->n[2] + 64 * a.vec - 1) / (64 * a.vec);
    a.nby = (s->n[1] + s->march_ty - 1) / s->march_ty;
    const long nbx = (s->n[0] + a.xchunk - 1) / a.xchunk;
    const long nb = (long)a.nbz * a.nby * nbx;
    sp.s = s->dev;
    sp.ops = s->ops;
    sp.us = usD;
    sp.usR = usR;
    sp.u0 = u0;
    sp.acc = acc;
    sp.out = out;
    sp.recs = s->special;
    sp.rows0 = s->srows[0];
    sp.rows1 = s->srows[1];
    sp.rows2 = s->srows[2];
    sp.nspecial = s->nspecial;
    sp.nR0 = s->nR[0];
    sp.nR1 = s->nR[1];
    sp.nR2 = s->nR[2];
    sp.pos0 = s->rpos[0];
    sp.pos1 = s->rpos[1];
    sp.pos2 = s->rpos[2];
    sp.dir0 = s->rdir[0];
    sp.dir1 = s->rdir[1];
    sp.dir2 = s->rdir[2];
    sp.outR0 = outR[0];
    sp.outR1 = outR[1];
    sp.outR2 = outR[2];
    sp.nfree = s->fused_free ? s->nfree : 0;
    sp.frees = s->frees;
    sp.fents = s->fents;
    sp.nface = s->nface_items;
    sp.u0R = u0R;
    sp.accF = accF;
    nitems = sp.nspecial + sp.nface + sp.nR0 + sp.nR1 + sp.nR2 + sp.nfree;
    if (nb + nitems / 64 + 8 >= (1L << 31)) return fail(s->ctx, SHX_EUNSUPPORTED, "march: too many workgroups");
    a.nblocks = (unsigned)nb;
    return SHX_OK;
}

// Stage register pointers of stage `stage` of a fused step on u0: input,
// output (FP64 or FP32 stage buffers, u0 at the ends) and cut points.
// The fused form's registers: the accumulator, the stage ping-pong (FP64,
// or FP32 for the mixed form) and the stage's cut-point values (FP64).  They
// are the system's own buffers, or a carve of the per-device workspace for
// the allocation-free entry point.
struct FusedRegs {
    double* acc = nullptr;
    double* accF = nullptr;  // accumulator of the free cut points
    double* stageD[2] = {nullptr, nullptr};
    float* stageF[2] = {nullptr, nullptr};
    double* stageR[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    // when keep[0] is set (FP64): stage s < 3 writes keep[s], which stage s + 1
    // reads, and stage 3 writes out; the step's input is only read
    shx_field* keep[3] = {nullptr, nullptr, nullptr};
    shx_field* out = nullptr;
};

FusedRegs own_regs(const shx_system* s)
{
    FusedRegs r;
    r.acc = s->accD;
    r.accF = s->accF;
    for (int b = 0; b < 2; ++b) {
        r.stageD[b] = s->stageD[b];
        r.stageF[b] = s->stageF[b];
        for (int c = 0; c < 3; ++c) r.stageR[b][c] = s->stageR[b][c];
    }
    return r;
}

void stage_registers(shx_system* s, const FusedRegs& g, shx_field* u0, int stage, const void*& usV,
                     void*& outV, RPtrs& usR, double* outR[3])
{
    if (g.keep[0]) {
        usV = stage == 0 ? (const void*)u0->D : (const void*)g.keep[stage - 1]->D;
        outV = stage == 3 ? (void*)g.out->D : (void*)g.keep[stage]->D;
        for (int c = 0; c < 3; ++c) {
            usR.r[c] = stage == 0 ? u0->R[c] : g.keep[stage - 1]->R[c];
            outR[c] = stage == 3 ? g.out->R[c] : g.keep[stage]->R[c];
        }
        return;
    }
    const bool mixed = s->precision == SHX_F32_COMP;
    usV = stage == 0 ? (const void*)u0->D
                     : mixed ? (const void*)g.stageF[(stage - 1) & 1] : (const void*)g.stageD[(stage - 1) & 1];
    outV = stage == 3 ? (void*)u0->D : mixed ? (void*)g.stageF[stage & 1] : (void*)g.stageD[stage & 1];
    for (int c = 0; c < 3; ++c) {
        usR.r[c] = stage == 0 ? u0->R[c] : g.stageR[(stage - 1) & 1][c];
        outR[c] = stage == 3 ? u0->R[c] : g.stageR[stage & 1][c];
    }
}

// One stage of the x-marching form in one launch: march tiles for interior,
// solid and pattern points; special items for the listed points and the
// Dirichlet cut points.
int launch_march(shx_system* s, int stage, const void* usD, RPtrs usR, const double* u0, RPtrs u0R,
                 double* acc, double* accF, void* out, double* const outR[3], double ts, double c_acc,
                 double c_next, double tn)
{
    MarchArgs a;
    SpecialArgs sp;
    long nitems = 0;
    int rc = march_args(s, usD, usR, u0, u0R, acc, accF, out, outR, a, sp, nitems);
    if (rc != SHX_OK) return rc;
    const StageScalars ss{ts, c_acc, c_next, tn};
    const int lrc = s->precision == SHX_F32_COMP
                        ? launch_stage_march_mixed(s->ctx->stream, s->family, stage, s->width,
                                                   s->march_ty, s->variant - 1, a, sp, ss)
                        : launch_stage_march(s->ctx->stream, s->family, stage, s->width, s->march_ty,
                                             s->variant - 1, a, sp, ss);
    if (lrc != 0)
        return fail(s->ctx, SHX_EINVAL, "march: unsupported tile height or stencil width");
    return check(s->ctx, "rk4 march stage kernel");
}

// One fused step on the registers g; when ev is given, event pairs bracket
// each stage's kernels (the point-wise tiles, or the x-marching launch or
// launches; the point-wise form's cut-point update runs outside the
// brackets).  The registers need no initial contents: stage 0 writes the
// accumulator and the first stage register before any stage reads them.
int rk4_fused_step_on(shx_system* s, const FusedRegs& g, shx_field* u0, double t, double dt,
                      hipEvent_t* ev)
{
    const bool march = s->variant >= 1;
    int rc = SHX_OK;
    for (int stage = 0; stage < 4 && rc == SHX_OK; ++stage) {
        const void* usV;
        void* outV;
        RPtrs usR;
        double* outR[3];
        stage_registers(s, g, u0, stage, usV, outV, usR, outR);
        const double* usD = static_cast<const double*>(usV);
        double* outD = static_cast<double*>(outV);
        const double ts = t + dt * kRki[stage];
        const double c_acc = dt * kRkf[stage];
        const double c_next = stage < 3 ? dt * kRki[stage + 1] : 1.0;
        const double tn = stage < 3 ? t + dt * kRki[stage + 1] : t + dt;
        if (ev) hipEventRecord(ev[2 * stage], s->ctx->stream);
        if (march)
            rc = launch_march(s, stage, usV, usR, u0->D, rptrs(u0), g.acc, g.accF, outV, outR, ts, c_acc,
                              c_next, tn);
        else if (s->family == SHX_SYS_HEAT)
            rc = launch_stage<SHX_SYS_HEAT>(s, stage, usD, usR, u0->D, g.acc, outD, ts, c_acc, c_next, tn);
        else if (s->family == kHeatGauss)
            rc = launch_stage<kHeatGauss>(s, stage, usD, usR, u0->D, g.acc, outD, ts, c_acc, c_next, tn);
        else if (s->kind == SHX_SYS_WAVE_RADIAL)
            rc = launch_stage<SHX_SYS_WAVE_RADIAL>(s, stage, usD, usR, u0->D, g.acc, outD, ts, c_acc, c_next, tn);
        else
            rc = launch_stage<SHX_SYS_WAVE_LINEAR>(s, stage, usD, usR, u0->D, g.acc, outD, ts, c_acc, c_next, tn);
        if (ev) hipEventRecord(ev[2 * stage + 1], s->ctx->stream);
        if (rc == SHX_OK && !march) rc = r_exact_all(s, outR, tn);
    }
    return rc;
}

// Whether the fused form can step this system at all.
int fused_supported(shx_system* s)
{
    if (s->hm)
        return fail(s->ctx, SHX_EUNSUPPORTED,
                    "fused form: a host-evaluated manufactured solution takes the reference schedule (shx_rk4_step)");
    if (s->general && !s->fused_free)
        return fail(s->ctx, SHX_EUNSUPPORTED,
                    "fused form: Neumann grid faces take the reference schedule (shx_rk4_step)");
    if (s->general && s->variant < 1)
        return fail(s->ctx, SHX_EUNSUPPORTED,
                    "fused form: floating/Neumann bodies need the x-marching variants (1-4)");
    if (s->precision == SHX_F32_COMP && s->variant < 1)
        return fail(s->ctx, SHX_EUNSUPPORTED, "mixed precision needs the x-marching form (variants 1-4)");
    return SHX_OK;
}

// One fused step on the system's own registers (allocated on first use,
// with the x-marching form's point classes).
int rk4_fused_step(shx_system* s, shx_field* u0, double t, double dt, hipEvent_t* ev)
{
    int rc = fused_supported(s);
    if (rc == SHX_OK) rc = ensure_fused(s);
    if (rc == SHX_OK && s->variant >= 1) rc = ensure_classes(s);
    if (rc == SHX_OK) rc = rk4_fused_step_on(s, own_regs(s), u0, t, dt, ev);
    return rc;
}

// One fused FP64 step from u (read only) to out, keeping the stage inputs
// w1, w2, w3 (the adjoint's linearization points).
int rk4_fused_keep(shx_system* s, const shx_field* u, shx_field* w1, shx_field* w2, shx_field* w3, shx_field* out,
                   double t, double dt)
{
    if (s->precision != SHX_F64)
        return fail(s->ctx, SHX_EUNSUPPORTED, "kept stage inputs: FP64 systems only");
    int rc = fused_supported(s);
    if (rc == SHX_OK) rc = ensure_fused(s);
    if (rc == SHX_OK && s->variant >= 1) rc = ensure_classes(s);
    if (rc != SHX_OK) return rc;
    FusedRegs g = own_regs(s);
    g.keep[0] = w1;
    g.keep[1] = w2;
    g.keep[2] = w3;
    g.out = out;
    return rk4_fused_step_on(s, g, const_cast<shx_field*>(u), t, dt, nullptr);
}

// The fused registers carved from the per-device workspace, 256-byte aligned.
int carve_workspace(shx_system* s, FusedRegs& g)
{
    auto* base = static_cast<unsigned char*>(s->ctx->scratch);
    if (!base) return fail(s->ctx, SHX_EHIP, "no-alloc fused steps: the device workspace is not available");
    const bool mixed = s->precision == SHX_F32_COMP;
    const size_t N = (size_t)s->N;
    auto round = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t need = round(N * sizeof(double)) + 2 * round(N * (mixed ? sizeof(float) : sizeof(double)));
    for (int c = 0; c < 3; ++c) need += 2 * round((size_t)s->nR[c] * sizeof(double));
    need += round((size_t)s->nfree * sizeof(double));
    if (need > kScratchBytes)
        return fail(s->ctx, SHX_EINVAL,
                    "no-alloc fused steps: the registers of " + std::to_string(s->N) + " points need " +
                        std::to_string(need) + " bytes, more than the " + std::to_string(kScratchBytes) +
                        "-byte device workspace");
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void* p = base + off;
        off += round(bytes);
        return p;
    };
    g.acc = static_cast<double*>(take(N * sizeof(double)));
    for (int b = 0; b < 2; ++b) {
        if (mixed) g.stageF[b] = static_cast<float*>(take(N * sizeof(float)));
        else g.stageD[b] = static_cast<double*>(take(N * sizeof(double)));
    }
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 3; ++c) g.stageR[b][c] = static_cast<double*>(take((size_t)s->nR[c] * sizeof(double)));
    g.accF = static_cast<double*>(take((size_t)s->nfree * sizeof(double)));
    return SHX_OK;
}

}  // namespace

// ------------------------------------------------------------------------ C ABI
extern "C" {

int shx_system_create(shx_ctx* ctx, const shx_system_desc* desc, shx_system** out)
{
    if (!ctx || !desc || !out) return SHX_EINVAL;
    *out = nullptr;
    if (desc->kind < SHX_SYS_WAVE_RADIAL || desc->kind > SHX_SYS_HEAT)
        return fail(ctx, SHX_EINVAL, "system: unknown kind");
    bool free_r = false, neumann = false;
    for (int c = 0; c < 3; ++c)
        for (long i = 0; i < desc->nR[c]; ++i) free_r = free_r || !desc->r_dirichlet[c][i];
    for (int f = 0; f < 6; ++f) neumann = neumann || desc->grid_bc[f] == SHX_BC_NEUMANN;

    auto* s = new (std::nothrow) shx_system();
    if (!s) return SHX_ENOMEM;
    s->ctx = ctx;
    s->kind = desc->kind;
    s->family = desc->kind == SHX_SYS_HEAT && desc->mms_kind == SHX_MMS_GAUSS ? kHeatGauss : desc->kind;
    s->general = free_r || neumann;
    s->any_free_r = free_r;
    for (int c = 0; c < 3; ++c) {
        s->n[c] = desc->n[c];
        s->nR[c] = desc->nR[c];
        s->d[c] = desc->d[c];
    }
    s->N = (long)s->n[0] * s->n[1] * s->n[2];
    s->tiling = make_tiling(s->n);

    std::memset(&s->dev, 0, sizeof s->dev);
    DevSys& dv = s->dev;
    dv.kind = desc->kind;
    dv.n0 = s->n[0];
    dv.n1 = s->n[1];
    dv.n2 = s->n[2];
    for (int f = 0; f < 6; ++f) dv.bc[f] = desc->grid_bc[f];
    for (int c = 0; c < 3; ++c) {
        dv.center[c] = desc->center[c];
        dv.vel[c] = desc->vel[c];
        dv.x0[c] = desc->x0[c];
    }
    dv.radius = desc->radius;
    dv.width = desc->width;
    dv.amp = desc->amp;
    dv.k = desc->diffusivity;
    dv.mms_kind = desc->mms_kind;
    dv.mms_dims = desc->mms_dims;
    dv.mms_n = desc->mms_n;

    int rc = upload(ctx, &s->xyz[0], desc->x, (size_t)s->n[0]);
    if (rc == SHX_OK) rc = upload(ctx, &s->xyz[1], desc->y, (size_t)s->n[1]);
    if (rc == SHX_OK) rc = upload(ctx, &s->xyz[2], desc->z, (size_t)s->n[2]);
    if (rc == SHX_OK && desc->mms_n > 0) rc = upload(ctx, &s->mms, desc->mms, (size_t)8 * desc->mms_n);
    if (desc->kind == SHX_SYS_HEAT && desc->mms_kind == SHX_MMS_SINPROD)
        for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
            rc = alloc_d(ctx, &s->sintab[c], s->n[c]);
            if (rc == SHX_OK) {
                k_sin_table<<<(unsigned)((s->n[c] + 255) / 256), 256, 0, ctx->stream>>>(s->xyz[c], s->n[c],
                                                                                      s->sintab[c]);
                rc = check(ctx, "sin table kernel");
            }
        }
    // fluid gather and bit mask
    std::vector<unsigned> mask((s->N + 31) / 32, 0u);
    std::vector<int> fl(desc->nfluid);
    for (long q = 0; q < desc->nfluid; ++q) {
        const long idx = desc->fluid[q];
        mask[idx >> 5] |= 1u << (idx & 31);
        fl[q] = (int)idx;
    }
    s->nfluid = desc->nfluid;
    if (rc == SHX_OK) rc = upload(ctx, &s->fluid_mask, mask.data(), mask.size());
    if (rc == SHX_OK) rc = upload(ctx, &s->fluid, fl.data(), fl.size());
    // Dirichlet face list in for_each_grid_bc_desc order (duplicates kept)
    std::vector<int> faces;
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    for (int dir = 0; dir < 3; ++dir)
        for (int side = 0; side < 2; ++side) {
            if (desc->grid_bc[2 * dir + side] != SHX_BC_DIRICHLET) continue;
            const int p = side ? s->n[dir] - 1 : 0;
            if (dir == 0)
                for (long q = 0; q < (long)ny * nz; ++q) faces.push_back((int)((long)p * ny * nz + q));
            else if (dir == 1)
                for (int i = 0; i < nx; ++i)
                    for (int k = 0; k < nz; ++k) faces.push_back((int)((long)i * ny * nz + (long)p * nz + k));
            else
                for (long q = 0; q < (long)nx * ny; ++q) faces.push_back((int)(p + q * nz));
        }
    s->nfaces = (long)faces.size();
    if (rc == SHX_OK) rc = upload(ctx, &s->faces, faces.data(), faces.size());
    // heat with Neumann faces: the points of each Neumann face and the data
    // field N nu reads (zero until update_boundary fills it)
    if (rc == SHX_OK && neumann && desc->kind == SHX_SYS_HEAT) {
        std::vector<int> nf;
        for (int f = 0; f < 6; ++f) {
            s->neu_off[f] = (long)nf.size();
            if (desc->grid_bc[f] != SHX_BC_NEUMANN) continue;
            const int dir = f / 2, p = f % 2 ? s->n[dir] - 1 : 0;
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int k = 0; k < nz; ++k) {
                        const int q[3] = {i, j, k};
                        if (q[dir] == p) nf.push_back((int)(((long)i * ny + j) * nz + k));
                    }
        }
        s->neu_off[6] = (long)nf.size();
        rc = upload(ctx, &s->neu_faces, nf.data(), nf.size());
        if (rc == SHX_OK) rc = alloc_d(ctx, &s->neuD, s->N);
        if (rc == SHX_OK) rc = launch_fill(ctx, s->neuD, 0.0, s->N);
    }
    // SHX_MMS_HOST: host copies of every evaluation site, device staging
    if (rc == SHX_OK && desc->kind == SHX_SYS_HEAT && desc->mms_kind == SHX_MMS_HOST) {
        const shx_mms_fns* f = desc->mms_fns;
        if (!f || !f->call || !f->ddt || !f->grad || !f->lap) {
            rc = fail(ctx, SHX_EINVAL, "system: a host-evaluated manufactured solution needs call, ddt, grad and lap");
        } else {
            s->hm = new (std::nothrow) HostMms();
            if (!s->hm) rc = SHX_ENOMEM;
        }
        if (rc == SHX_OK) {
            HostMms& h = *s->hm;
            h.f = *f;
            h.xyz[0].assign(desc->x, desc->x + s->n[0]);
            h.xyz[1].assign(desc->y, desc->y + s->n[1]);
            h.xyz[2].assign(desc->z, desc->z + s->n[2]);
            long most = s->N;
            for (int c = 0; c < 3; ++c) {
                h.pos[c].assign(desc->r_pos[c], desc->r_pos[c] + 3 * s->nR[c]);
                h.dirich[c].resize((size_t)s->nR[c]);
                for (long i = 0; i < s->nR[c]; ++i) h.dirich[c][(size_t)i] = desc->r_dirichlet[c][i] ? 1 : 0;
                most = std::max(most, s->nR[c]);
            }
            h.fluid = fl;
            h.faces = faces;
            if (s->neu_faces) {
                for (int f6 = 0; f6 < 7; ++f6) h.neu_off[f6] = s->neu_off[f6];
                h.neu.resize((size_t)s->neu_off[6]);
                const hipError_t e = hipMemcpy(h.neu.data(), s->neu_faces, h.neu.size() * sizeof(int), hipMemcpyDeviceToHost);
                if (e != hipSuccess) rc = hip_fail(ctx, e, "neumann faces d2h");
            }
            if (rc == SHX_OK) rc = alloc_d(ctx, &h.d, most);
            if (rc == SHX_OK) rc = alloc_d(ctx, &h.exD, s->nfluid);
        }
    }
    for (int c = 0; c < 3 && rc == SHX_OK; ++c)
        if (free_r && s->nR[c] > 0) rc = alloc_d(ctx, &s->exR[c], s->nR[c]);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
        rc = upload(ctx, &s->rpos[c], desc->r_pos[c], (size_t)(3 * s->nR[c]));
        std::vector<signed char> fl8(s->nR[c]);
        for (long i = 0; i < s->nR[c]; ++i) fl8[i] = desc->r_dirichlet[c][i] ? 1 : 0;
        if (rc == SHX_OK)
            rc = upload(ctx, reinterpret_cast<int8_t**>(&s->rdir[c]),
                        reinterpret_cast<const int8_t*>(fl8.data()), fl8.size());
    }
    // the fused form's free cut points (no Neumann grid faces): point q of
    // R(r) and the entries of its row in each derivative's R-space row list
    // (rows sorted by (set, row)), copied in direction order: the first
    // kFreeSlots into the point's slots, the rest after every point's slots
    if (rc == SHX_OK && free_r && !neumann) {
        std::vector<FreeRec> fr;
        std::vector<std::vector<FreeEnt>> per;
        for (int r = 0; r < 3; ++r)
            for (long q = 0; q < s->nR[r]; ++q) {
                if (desc->r_dirichlet[r][q]) continue;
                FreeRec e{};
                e.q = q;
                e.set = r;
                std::vector<FreeEnt> ents;
                for (int c = 0; c < 3; ++c) {
                    e.n[c] = 0;
                    const shx_derivative* d = s->d[c];
                    if (!d || d->h_rset.empty()) continue;
                    long lo = 0, hi = (long)d->h_rset.size();
                    while (lo < hi) {  // first row not before (r, q)
                        const long mid = (lo + hi) / 2;
                        if (d->h_rset[mid] < r || (d->h_rset[mid] == r && d->h_rrow[mid] < q)) lo = mid + 1;
                        else hi = mid;
                    }
                    if (lo >= (long)d->h_rset.size() || d->h_rset[lo] != r || d->h_rrow[lo] != q) continue;
                    for (long k = d->h_rptr[(size_t)lo]; k < d->h_rptr[(size_t)lo + 1]; ++k)
                        ents.push_back(FreeEnt{d->h_rcol[(size_t)k], d->h_rw[(size_t)k]});
                    e.n[c] = (int)(d->h_rptr[(size_t)lo + 1] - d->h_rptr[(size_t)lo]);
                }
                fr.push_back(e);
                per.push_back(std::move(ents));
            }
        s->nfree = (long)fr.size();
        std::vector<FreeEnt> list((size_t)s->nfree * kFreeSlots, FreeEnt{0, 0.0});
        for (size_t f = 0; f < per.size(); ++f) {
            fr[f].over = (long)list.size();
            for (size_t m = 0; m < per[f].size(); ++m) {
                if (m < (size_t)kFreeSlots) list[f * kFreeSlots + m] = per[f][m];
                else list.push_back(per[f][m]);
            }
        }
        rc = upload_vec(ctx, &s->frees, fr);
        if (rc == SHX_OK && !list.empty()) rc = upload_vec(ctx, &s->fents, list);
        s->fused_free = rc == SHX_OK;
    }
    if (rc == SHX_OK) {
        const long nparts = 1024;
        hipError_t e = hipMalloc((void**)&s->part, (size_t)nparts * 3 * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void**)&s->part_idx, (size_t)nparts * sizeof(long));
        if (e != hipSuccess) rc = hip_fail(ctx, e, "hipMalloc");
    }
    if (rc != SHX_OK) {
        shx_system_destroy(s);
        return rc;
    }
    dv.x = s->xyz[0];
    dv.y = s->xyz[1];
    dv.z = s->xyz[2];
    dv.sinx = s->sintab[0];
    dv.siny = s->sintab[1];
    dv.sinz = s->sintab[2];
    dv.mms = s->mms;
    dv.fluid_mask = s->fluid_mask;
    s->width = 0;
    for (int c = 0; c < 3; ++c) {
        if (s->d[c]) s->ops.d[c] = s->d[c]->dev;
        else std::memset(&s->ops.d[c], 0, sizeof(DevDeriv));
        if (s->ops.d[c].present) {
            if (s->width && s->width != s->ops.d[c].width) {
                shx_system_destroy(s);
                return fail(ctx, SHX_EINVAL, "system: derivatives with different stencil widths");
            }
            s->width = s->ops.d[c].width;
        }
    }
    if (s->width == 0) s->width = 5;
    // the x-marching form, special workgroups after the tiles, is the default
    // where it applies
    s->variant = march_ok(s) ? 3 : 0;
    *out = s;
    return SHX_OK;
}

int shx_system_destroy(shx_system* s)
{
    if (!s) return SHX_EINVAL;
    hipStreamSynchronize(s->ctx->stream);
    for (int c = 0; c < 3; ++c) {
        hipFree(s->xyz[c]);
        hipFree(s->sintab[c]);
        hipFree(s->rpos[c]);
        hipFree(s->rdir[c]);
        hipFree(s->exR[c]);
        for (int b = 0; b < 2; ++b) hipFree(s->stageR[b][c]);
        for (int r = 0; r < 4; ++r) {
            hipFree(s->gG[c][r]);
            hipFree(s->du[c][r]);
        }
        if (s->Bfull[c]) shx_csr_destroy(s->Bfull[c]);
    }
    destroy_adjoint_ws(s->adj);
    if (s->hm) {
        hipFree(s->hm->d);
        hipFree(s->hm->exD);
        delete s->hm;
    }
    hipFree(s->stageD[0]);
    hipFree(s->stageD[1]);
    hipFree(s->stageF[0]);
    hipFree(s->stageF[1]);
    hipFree(s->accD);
    hipFree(s->accF);
    hipFree(s->frees);
    hipFree(s->fents);
    hipFree(s->cls);
    hipFree(s->special);
    for (int c = 0; c < 3; ++c) hipFree(s->srows[c]);
    hipFree(s->crows);
    hipFree(s->neuD);
    hipFree(s->neu_faces);
    hipFree(s->srcD);
    hipFree(s->mms);
    hipFree(s->fluid_mask);
    hipFree(s->fluid);
    hipFree(s->faces);
    hipFree(s->part);
    hipFree(s->part_idx);
    delete s;
    return SHX_OK;
}

int shx_system_rhs(shx_system* s, const shx_field* u, shx_field* rhs, double t)
{
    if (!s || !u || !rhs) return SHX_EINVAL;
    int rc = rhs_fused_launch(s, u, rhs, t);
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_system_rhs_unfused(shx_system* s, const shx_field* u, shx_field* rhs, double t)
{
    if (!s || !u || !rhs) return SHX_EINVAL;
    int rc = rhs_unfused_launch(s, u, rhs, t);
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_system_update_boundary(shx_system* s, shx_field* u, double t)
{
    if (!s || !u) return SHX_EINVAL;
    int rc = update_boundary_launch(s, u, t);
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_system_exact(shx_system* s, shx_field* out, double t)
{
    if (!s || !out) return SHX_EINVAL;
    hipStream_t st = s->ctx->stream;
    if (s->hm) {
        hm_mesh(s, kHmCall, t, nullptr, s->N);
        int rc = hm_put(s, out->D);
        for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
            hm_cut(s, c, kHmCall, t);
            rc = hm_put(s, out->R[c]);
        }
        return rc;
    }
    k_exact_D<<<grid_for(s->N), 256, 0, st>>>(s->dev, s->N, out->D, t);
    for (int c = 0; c < 3; ++c)
        if (s->nR[c] > 0) k_exact_R<<<grid_for(s->nR[c]), 256, 0, st>>>(s->dev, s->nR[c], s->rpos[c], out->R[c], t);
    int rc = check(s->ctx, "exact kernel");
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_system_initialize(shx_system* s, shx_field* u, double t)
{
    if (!s || !u) return SHX_EINVAL;
    hipStream_t st = s->ctx->stream;
    int rc = launch_fill(s->ctx, u->D, 0.0, s->N);
    if (s->hm) {
        if (rc == SHX_OK && s->nfluid) {
            hm_mesh(s, kHmCall, t, s->hm->fluid.data(), s->nfluid);
            rc = hm_put(s, s->hm->d);
            if (rc == SHX_OK) k_scatter<<<grid_for(s->nfluid), 256, 0, st>>>(s->nfluid, s->fluid, s->hm->d, u->D);
            if (rc == SHX_OK) rc = check(s->ctx, "initialize kernels");
            if (rc == SHX_OK) rc = sync(s->ctx);
        }
        for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
            hm_cut(s, c, kHmCall, t);
            rc = hm_put(s, u->R[c]);
        }
        return rc;
    }
    if (rc == SHX_OK && s->nfluid)
        k_init_fluid<<<grid_for(s->nfluid), 256, 0, st>>>(s->dev, s->nfluid, s->fluid, u->D, t);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c)
        if (s->nR[c] > 0) k_exact_R<<<grid_for(s->nR[c]), 256, 0, st>>>(s->dev, s->nR[c], s->rpos[c], u->R[c], t);
    if (rc == SHX_OK) rc = check(s->ctx, "initialize kernels");
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_system_stats(shx_system* s, const shx_field* u, double t, double stats[11])
{
    if (!s || !u || !stats) return SHX_EINVAL;
    const long nparts = 1024;
    double* pe = s->part;
    double* pmin = s->part + nparts;
    double* pmax = s->part + 2 * nparts;
    for (int q = 0; q < 11; ++q) stats[q] = 0.0;
    if (s->nfluid > 0) {
        const unsigned grid = (unsigned)std::min<long>(nparts, (s->nfluid + 255) / 256);
        const double* exF = nullptr;
        if (s->hm) {  // the exact values on the fluid gather, from the host
            hm_mesh(s, kHmCall, t, s->hm->fluid.data(), s->nfluid);
            const int rc0 = hm_put(s, s->hm->exD);
            if (rc0 != SHX_OK) return rc0;
            exF = s->hm->exD;
        }
        k_stats<<<grid, 256, 0, s->ctx->stream>>>(s->dev, s->nfluid, s->fluid, u->D, exF, t, pe, s->part_idx,
                                                   pmin, pmax);
        int rc = check(s->ctx, "stats kernel");
        if (rc != SHX_OK) return rc;
        std::vector<double> h(3 * nparts);
        std::vector<long> hi(nparts);
        hipError_t e = hipMemcpyAsync(h.data(), s->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hi.data(), s->part_idx, hi.size() * sizeof(long), hipMemcpyDeviceToHost, s->ctx->stream);
        if (e != hipSuccess) return hip_fail(s->ctx, e, "stats copy");
        rc = sync(s->ctx);
        if (rc != SHX_OK) return rc;
        // MaxLoc identity (lowest value, largest index) as Kokkos 5.1's
        // MaxLoc::init, so that a field without a strict maximum (all NaN)
        // reports what the reference reports
        double err = -DBL_MAX, mn = DBL_MAX, mx = -DBL_MAX;
        long ei = 2147483647L;
        for (unsigned b = 0; b < grid; ++b) {
            if (h[b] > err || (h[b] == err && hi[b] < ei)) {
                err = h[b];
                ei = hi[b];
            }
            mn = std::min(mn, h[nparts + b]);
            mx = std::max(mx, h[2 * nparts + b]);
        }
        // err = max(err_d, err_rx, err_ry, err_rz) with the R terms 0 for
        // Dirichlet bodies (scalar_system_utils.hpp:169)
        stats[0] = err < 0.0 ? 0.0 : err;
        stats[1] = mn;
        stats[2] = mx;
        stats[3] = err;
        stats[4] = (double)ei;
    }
    // the cut points of non-Dirichlet bodies (scalar_system_utils.hpp:120-169):
    // their values join min and max, their errors give Rx/Ry/Rz Linf and
    // index (first strict maximum), and the overall Linf; few points, so on
    // the host
    if (s->any_free_r) {
        if (s->nfluid == 0) stats[1] = stats[2] = stats[3] = 0.0;
        double mn = s->nfluid ? stats[1] : 0.0, mx = s->nfluid ? stats[2] : 0.0;
        double err = stats[3];
        for (int c = 0; c < 3; ++c) {
            const long n = s->nR[c];
            if (n == 0) continue;
            int rc = SHX_OK;
            if (s->hm) {
                hm_cut(s, c, kHmCall, t);
                rc = hm_put(s, s->exR[c]);
                if (rc != SHX_OK) return rc;
            } else {
                k_exact_R<<<grid_for(n), 256, 0, s->ctx->stream>>>(s->dev, n, s->rpos[c], s->exR[c], t);
                rc = check(s->ctx, "exact kernel");
            }
            std::vector<double> hu(n), hx(n);
            std::vector<signed char> fl(n);
            hipError_t e = hipMemcpyAsync(hu.data(), u->R[c], n * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(hx.data(), s->exR[c], n * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(fl.data(), s->rdir[c], n, hipMemcpyDeviceToHost, s->ctx->stream);
            if (e != hipSuccess) return hip_fail(s->ctx, e, "stats copy");
            if (rc == SHX_OK) rc = sync(s->ctx);
            if (rc != SHX_OK) return rc;
            double best = -DBL_MAX;
            long bi = 2147483647L;
            bool any = false;
            for (long i = 0; i < n; ++i) {
                if (fl[i]) continue;
                any = true;
                if (hu[i] < mn) mn = hu[i];
                if (hu[i] > mx) mx = hu[i];
                const double ev = std::abs(hu[i] - hx[i]);
                if (ev > best) {
                    best = ev;
                    bi = i;
                }
            }
            if (any) {
                stats[5 + 2 * c] = best;
                stats[6 + 2 * c] = (double)bi;
                if (best > err) err = best;
            }
        }
        stats[0] = err;
        stats[1] = mn;
        stats[2] = mx;
    }
    return SHX_OK;
}

int shx_euler_step(shx_system* s, shx_field* u0, shx_field* out, shx_field* srhs, double t,
                   double dt)
{
    if (!s || !u0 || !out || !srhs) return SHX_EINVAL;
    int rc = slot_copy(s, out, u0);
    if (rc == SHX_OK) rc = zero_slot(s, srhs);
    if (rc == SHX_OK) rc = rhs_unfused_launch(s, out, srhs, t);
    if (rc == SHX_OK) rc = slot_assign_lc(s, out, u0, dt, srhs);
    if (rc == SHX_OK) rc = update_boundary_launch(s, out, t + dt);
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_rk4_step(shx_system* s, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs,
                 double t, double dt)
{
    if (!s || !u0 || !out || !rk || !srhs) return SHX_EINVAL;
    int rc = rk4_unfused(s, u0, out, rk, srhs, t, dt);
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_rk4_fused_keep_async(shx_system* s, const shx_field* u, shx_field* w1, shx_field* w2, shx_field* w3,
                             shx_field* out, double t, double dt)
{
    if (!s || !u || !w1 || !w2 || !w3 || !out || out == u) return SHX_EINVAL;
    return rk4_fused_keep(s, u, w1, w2, w3, out, t, dt);
}

int shx_rk4_fused_steps_async(shx_system* s, shx_field* u0, double t, double dt, int nsteps)
{
    if (!s || !u0 || nsteps < 0) return SHX_EINVAL;
    int rc = SHX_OK;
    for (int n = 0; n < nsteps && rc == SHX_OK; ++n) rc = rk4_fused_step(s, u0, t + n * dt, dt, nullptr);
    return rc;
}

int shx_rk4_fused_steps_noalloc(shx_system* s, shx_field* u0, double t, double dt, int nsteps)
{
    if (!s || !u0 || nsteps < 0) return SHX_EINVAL;
    int rc = fused_supported(s);
    if (rc == SHX_OK && s->variant >= 1 && !s->classes_ready)
        return fail(s->ctx, SHX_EINVAL,
                    "no-alloc fused steps: the point classes of the x-marching form are not built "
                    "(call shx_system_stage_points, or take one fused step, once first)");
    FusedRegs g;
    if (rc == SHX_OK) rc = carve_workspace(s, g);
    for (int n = 0; n < nsteps && rc == SHX_OK; ++n) rc = rk4_fused_step_on(s, g, u0, t + n * dt, dt, nullptr);
    return rc;
}

int shx_time_fused_steps(shx_system* s, shx_field* u0, double t, double dt, int nsteps, int reps,
                         double* ms_per_step)
{
    if (!s || !u0 || nsteps < 1 || reps < 1 || !ms_per_step) return SHX_EINVAL;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    int rc = ensure_fused(s);
    if (rc == SHX_OK) rc = sync(s->ctx);
    hipEventRecord(a, s->ctx->stream);
    for (int r = 0; r < reps && rc == SHX_OK; ++r)
        for (int n = 0; n < nsteps && rc == SHX_OK; ++n) rc = rk4_fused_step(s, u0, t + n * dt, dt, nullptr);
    hipEventRecord(b, s->ctx->stream);
    if (rc == SHX_OK) rc = sync(s->ctx);
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    *ms_per_step = (double)ms / ((double)reps * nsteps);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return rc;
}

int shx_time_unfused_steps(shx_system* s, shx_field* u0, shx_field* out, shx_field* rk,
                           shx_field* srhs, double t, double dt, int nsteps, double* ms_per_step)
{
    if (!s || !u0 || !out || !rk || !srhs || nsteps < 1 || !ms_per_step) return SHX_EINVAL;
    int rc = ensure_unfused(s);
    if (rc != SHX_OK) return rc;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, s->ctx->stream);
    for (int n = 0; n < nsteps && rc == SHX_OK; ++n) {
        rc = rk4_unfused(s, u0, out, rk, srhs, t + n * dt, dt);
        // simulation_cycle.cpp:122-125: u0 <- u1 by copy
        if (rc == SHX_OK) rc = slot_copy(s, u0, out);
    }
    hipEventRecord(b, s->ctx->stream);
    if (rc == SHX_OK) rc = sync(s->ctx);
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    *ms_per_step = (double)ms / nsteps;
    hipEventDestroy(a);
    hipEventDestroy(b);
    return rc;
}

// Mean device time of one fused stage kernel, bracketed by HIP events on the
// stream it is launched on (the cut-point kernels are outside the brackets).
int shx_time_fused_stage_kernel(shx_system* s, shx_field* u0, double t, double dt, int nsteps,
                                double* ms_per_kernel)
{
    if (!s || !u0 || nsteps < 1 || !ms_per_kernel) return SHX_EINVAL;
    int rc = ensure_fused(s);
    if (rc == SHX_OK) rc = sync(s->ctx);
    std::vector<hipEvent_t> ev(8 * (size_t)nsteps);
    for (auto& e : ev) hipEventCreate(&e);
    for (int n = 0; n < nsteps && rc == SHX_OK; ++n) rc = rk4_fused_step(s, u0, t + n * dt, dt, &ev[8 * (size_t)n]);
    if (rc == SHX_OK) rc = sync(s->ctx);
    double total = 0.0;
    for (int q = 0; q < 4 * nsteps; ++q) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]);
        total += ms;
    }
    *ms_per_kernel = total / (4.0 * nsteps);
    for (auto& e : ev) hipEventDestroy(e);
    return rc;
}

int shx_system_set_tuning(shx_system* s, int variant, int march_ty, int march_xchunk)
{
    if (!s || variant < 0 || variant > 4) return SHX_EINVAL;
    if (variant >= 1 && !(march_ty == 4 || march_ty == 8)) return SHX_EINVAL;
    if (variant >= 1 && march_xchunk < 1) return SHX_EINVAL;
    if (variant >= 1 && !march_ok(s)) return fail(s->ctx, SHX_EUNSUPPORTED, "x-marching form: unsupported system");
    if (variant == 0 && s->precision == SHX_F32_COMP)
        return fail(s->ctx, SHX_EUNSUPPORTED, "mixed precision needs the x-marching form (variants 1-4)");
    s->variant = variant;
    if (variant >= 1) {
        s->march_ty = march_ty;
        s->march_xchunk = march_xchunk;
    }
    return SHX_OK;
}

int shx_system_set_precision(shx_system* s, int precision)
{
    if (!s || (precision != SHX_F64 && precision != SHX_F32_COMP)) return SHX_EINVAL;
    if (precision == SHX_F32_COMP && !march_ok(s))
        return fail(s->ctx, SHX_EUNSUPPORTED, "mixed precision: unsupported system");
    if (precision == SHX_F32_COMP && s->variant == 0) s->variant = 3;
    s->precision = precision;
    return SHX_OK;
}

int shx_system_get_precision(const shx_system* s, int* precision)
{
    if (!s || !precision) return SHX_EINVAL;
    *precision = s->precision;
    return SHX_OK;
}

int shx_system_stage_points(shx_system* s, int64_t out[2])
{
    if (!s || !out) return SHX_EINVAL;
    if (s->variant >= 1) {
        const int rc = ensure_classes(s);
        if (rc != SHX_OK) return rc;
        out[0] = s->N;
        out[1] = s->nspecial + s->nface_points;
    } else {
        out[0] = s->N;
        out[1] = 0;
    }
    return SHX_OK;
}

int shx_system_get_tuning(const shx_system* s, int out[3])
{
    if (!s || !out) return SHX_EINVAL;
    out[0] = s->variant;
    out[1] = s->march_ty;
    out[2] = s->march_xchunk;
    return SHX_OK;
}

int shx_system_autotune(shx_system* s, const shx_field* u0, double t, double dt, double* best_ms)
{
    if (!s || !u0) return SHX_EINVAL;
    int rc = ensure_fused(s);
    if (rc != SHX_OK) return rc;
    // work on a copy so that the caller's state is untouched
    shx_field tmp;
    std::memset(&tmp, 0, sizeof tmp);
    rc = alloc_d(s->ctx, &tmp.D, s->N);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) rc = alloc_d(s->ctx, &tmp.R[c], s->nR[c]);
    auto reset = [&]() {
        hipMemcpyAsync(tmp.D, u0->D, sizeof(double) * s->N, hipMemcpyDeviceToDevice, s->ctx->stream);
        for (int c = 0; c < 3; ++c)
            if (s->nR[c])
                hipMemcpyAsync(tmp.R[c], u0->R[c], sizeof(double) * s->nR[c], hipMemcpyDeviceToDevice,
                               s->ctx->stream);
    };
    struct cand {
        int v, ty, xc;
    };
    std::vector<cand> cands;
    // the point-wise kernel steps neither the mixed form nor free cut points
    if (s->precision == SHX_F64 && !s->general) cands.push_back({0, 4, 32});
    if (march_ok(s))
        for (int v = 1; v <= 4; ++v)
            for (int ty : {4, 8})
                // chunk lengths that leave the tiles' last round of
                // workgroups partly free for the special work, as well as
                // the ones that fill whole rounds
                for (int xc : {8, 12, 16, 20, 24, 32, 64, 1 << 16}) {
                    const cand c{v, ty, std::min(xc, s->n[0])};
                    bool dup = false;
                    for (const cand& o : cands) dup = dup || (o.v == c.v && o.ty == c.ty && o.xc == c.xc);
                    if (!dup) cands.push_back(c);
                }
    if (cands.empty()) {
        for (int c = 0; c < 3; ++c) hipFree(tmp.R[c]);
        hipFree(tmp.D);
        return fail(s->ctx, SHX_EUNSUPPORTED, "autotune: no fused-stage kernel applies to this system");
    }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    double best = 1e30;
    cand pick = cands[0];
    for (const cand& c : cands) {
        if (rc != SHX_OK) break;
        s->variant = c.v;
        s->march_ty = c.ty;
        s->march_xchunk = c.xc;
        reset();
        rc = rk4_fused_step(s, &tmp, t, dt, nullptr);  // warm-up
        const int reps = 3;
        hipEventRecord(e0, s->ctx->stream);
        for (int r = 0; r < reps && rc == SHX_OK; ++r) rc = rk4_fused_step(s, &tmp, t + r * dt, dt, nullptr);
        hipEventRecord(e1, s->ctx->stream);
        if (rc == SHX_OK) rc = sync(s->ctx);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (rc == SHX_OK && ms / reps < best) {
            best = ms / reps;
            pick = c;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    s->variant = pick.v;
    s->march_ty = pick.ty;
    s->march_xchunk = pick.xc;
    hipFree(tmp.D);
    for (int c = 0; c < 3; ++c) hipFree(tmp.R[c]);
    if (best_ms) *best_ms = best;
    return rc;
}

// ------------------------------------------------------------------- batch
// Many problems stepped together: one launch per RK4 stage for all of them,
// the per-problem arguments and clocks in device memory.  No reference
// counterpart (the reference steps one problem per process).

}  // extern "C"

struct shx_batch {
    shx_ctx* ctx = nullptr;
    int n = 0;
    int kind = 0, width = 0, precision = SHX_F64, ty = 4, vec = 1;
    std::vector<shx_system*> sys;
    std::vector<shx_field> u;
    BatchEntry* ent[4] = {nullptr, nullptr, nullptr, nullptr};
    BatchClock* clk = nullptr;
    unsigned* start = nullptr;
    unsigned grid = 0;
    long points = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
};

namespace {

int batch_step(shx_batch* b)
{
    hipStream_t st = b->ctx->stream;
    for (int stage = 0; stage < 4; ++stage) {
        BatchLaunch bl{b->ent[stage], b->clk, b->start, b->n, b->vec, b->grid};
        const int lrc = b->precision == SHX_F32_COMP
                            ? launch_batch_stage_mixed(st, b->kind, stage, b->width, b->ty, bl)
                            : launch_batch_stage(st, b->kind, stage, b->width, b->ty, bl);
        if (lrc != 0) return fail(b->ctx, SHX_EINVAL, "batch: unsupported tile height or stencil width");
    }
    launch_batch_advance(st, b->clk, b->n);
    return check(b->ctx, "batch step");
}

void batch_free_graph(shx_batch* b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
}

}  // namespace

extern "C" {

int shx_batch_create(shx_ctx* ctx, shx_system* const* systems, const shx_field* u0, const double* t0,
                     const double* dt, int n, int march_ty, shx_batch** out)
{
    if (!ctx || !systems || !u0 || !t0 || !dt || n < 1 || !out) return SHX_EINVAL;
    if (!(march_ty == 4 || march_ty == 8)) return SHX_EINVAL;
    *out = nullptr;
    for (int p = 0; p < n; ++p)
        if (!systems[p] || (systems[p]->general && !systems[p]->fused_free) || systems[p]->hm)
            return fail(ctx, SHX_EUNSUPPORTED,
                        "batch: Neumann grid faces and host-evaluated solutions take the reference schedule");
    auto b = std::unique_ptr<shx_batch>(new (std::nothrow) shx_batch());
    if (!b) return SHX_ENOMEM;
    b->ctx = ctx;
    b->n = n;
    b->ty = march_ty;
    b->kind = systems[0]->family;
    b->width = systems[0]->width;
    b->precision = systems[0]->precision;
    // two points per lane along z when every problem's z extent is even
    b->vec = 2;
    for (int p = 0; p < n; ++p)
        if (systems[p] && systems[p]->n[2] % 2 != 0) b->vec = 1;
    std::vector<BatchEntry> ent[4];
    std::vector<BatchClock> clk;
    std::vector<unsigned> start;
    unsigned long long next = 0;
    for (int p = 0; p < n; ++p) {
        shx_system* s = systems[p];
        if (!s || s->ctx != ctx || s->family != b->kind || s->width != b->width || s->precision != b->precision)
            return fail(ctx, SHX_EINVAL, "batch: problems differ in context, system, stencil width or precision");
        if (!march_ok(s)) return fail(ctx, SHX_EUNSUPPORTED, "batch: the x-marching form does not apply");
        const int saved_ty = s->march_ty, saved_vec = s->march_vec;
        s->march_ty = march_ty;  // tile shape shared by the batch
        s->march_vec = b->vec;
        if (s->variant == 0) s->variant = 3;
        int rc = ensure_fused(s);
        if (rc == SHX_OK) rc = ensure_classes(s);
        if (rc != SHX_OK) {
            s->march_ty = saved_ty;
            s->march_vec = saved_vec;
            return rc;
        }
        b->sys.push_back(s);
        b->u.push_back(u0[p]);
        shx_field* up = &b->u.back();
        start.push_back((unsigned)next);
        unsigned sg = 0, mg = 0;
        for (int stage = 0; stage < 4; ++stage) {
            BatchEntry e{};
            const void* usV;
            void* outV;
            RPtrs usR;
            double* outR[3];
            stage_registers(s, own_regs(s), up, stage, usV, outV, usR, outR);
            long nitems = 0;
            rc = march_args(s, usV, usR, up->D, rptrs(up), s->accD, s->accF, outV, outR, e.m, e.sp, nitems);
            if (rc != SHX_OK) {
                s->march_ty = saved_ty;
                s->march_vec = saved_vec;
                return rc;
            }
            const long threads = 64L * march_ty;
            e.sg = sg = (unsigned)(((nitems + threads - 1) / threads + 7) / 8);
            e.mg = mg = (e.m.nblocks + 7u) / 8u;
            e.mode = 2;  // special work after the tiles
            ent[stage].push_back(e);
        }
        s->march_ty = saved_ty;
        s->march_vec = saved_vec;
        next += 8ull * (sg + mg);
        clk.push_back(BatchClock{t0[p], dt[p], 0});
        b->points += s->N;
    }
    if (next >= (1ull << 31)) return fail(ctx, SHX_EUNSUPPORTED, "batch: too many workgroups");
    b->grid = (unsigned)next;
    int rc = SHX_OK;
    for (int stage = 0; stage < 4 && rc == SHX_OK; ++stage) rc = upload_vec(ctx, &b->ent[stage], ent[stage]);
    if (rc == SHX_OK) rc = upload_vec(ctx, &b->clk, clk);
    if (rc == SHX_OK) rc = upload_vec(ctx, &b->start, start);
    if (rc != SHX_OK) {
        shx_batch_destroy(b.release());
        return rc;
    }
    *out = b.release();
    return SHX_OK;
}

int shx_batch_destroy(shx_batch* b)
{
    if (!b) return SHX_EINVAL;
    hipStreamSynchronize(b->ctx->stream);
    batch_free_graph(b);
    for (auto& e : b->ent) hipFree(e);
    hipFree(b->clk);
    hipFree(b->start);
    delete b;
    return SHX_OK;
}

int shx_batch_steps_async(shx_batch* b, int nsteps)
{
    if (!b || nsteps < 0) return SHX_EINVAL;
    int rc = SHX_OK;
    for (int k = 0; k < nsteps && rc == SHX_OK; ++k) rc = batch_step(b);
    return rc;
}

int shx_batch_capture(shx_batch* b, int nsteps)
{
    if (!b || nsteps < 1) return SHX_EINVAL;
    batch_free_graph(b);
    hipStream_t st = b->ctx->stream;
    int rc = sync(b->ctx);
    if (rc != SHX_OK) return rc;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(b->ctx, e, "batch capture");
    for (int k = 0; k < nsteps && rc == SHX_OK; ++k) rc = batch_step(b);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (rc != SHX_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(b->ctx, e, "batch capture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        batch_free_graph(b);
        return hip_fail(b->ctx, e, "batch graph instantiate");
    }
    b->graph_steps = nsteps;
    return SHX_OK;
}

int shx_batch_launch_async(shx_batch* b)
{
    if (!b) return SHX_EINVAL;
    if (!b->exec) return fail(b->ctx, SHX_EINVAL, "batch: no captured graph");
    const hipError_t e = hipGraphLaunch(b->exec, b->ctx->stream);
    return e == hipSuccess ? SHX_OK : hip_fail(b->ctx, e, "batch graph launch");
}

int shx_batch_times(shx_batch* b, double* t)
{
    if (!b || !t) return SHX_EINVAL;
    std::vector<BatchClock> clk((size_t)b->n);
    hipError_t e = hipMemcpyAsync(clk.data(), b->clk, sizeof(BatchClock) * clk.size(), hipMemcpyDeviceToHost,
                                  b->ctx->stream);
    if (e != hipSuccess) return hip_fail(b->ctx, e, "batch clocks");
    const int rc = sync(b->ctx);
    if (rc != SHX_OK) return rc;
    for (int p = 0; p < b->n; ++p) t[p] = clk[(size_t)p].t0 + (double)clk[(size_t)p].n * clk[(size_t)p].dt;
    return SHX_OK;
}

int shx_time_batch(shx_batch* b, int use_graph, int reps, double* ms_per_step, int64_t* points)
{
    if (!b || reps < 1 || !ms_per_step) return SHX_EINVAL;
    if (use_graph && !b->exec) return fail(b->ctx, SHX_EINVAL, "batch: no captured graph");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return hip_fail(b->ctx, hipGetLastError(), "event");
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return hip_fail(b->ctx, hipGetLastError(), "event");
    }
    int rc = sync(b->ctx);
    hipEventRecord(e0, b->ctx->stream);
    for (int r = 0; r < reps && rc == SHX_OK; ++r)
        rc = use_graph ? shx_batch_launch_async(b) : batch_step(b);
    hipEventRecord(e1, b->ctx->stream);
    if (rc == SHX_OK) rc = sync(b->ctx);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const int steps = use_graph ? b->graph_steps : 1;
    *ms_per_step = (double)ms / ((double)reps * steps);
    if (points) *points = b->points;
    return rc;
}

}  // extern "C"

// ---- one problem's fused RK4 steps as a hipGraph: a batch of that problem
// alone, captured (the clock lives on the device, so every replay advances)
struct shx_rk4_graph {
    shx_batch* b = nullptr;
};

extern "C" {

int shx_rk4_graph_create(shx_system* s, shx_field* u0, double t0, double dt, int nsteps, shx_rk4_graph** out)
{
    if (!s || !u0 || nsteps < 1 || !out) return SHX_EINVAL;
    *out = nullptr;
    auto g = std::unique_ptr<shx_rk4_graph>(new (std::nothrow) shx_rk4_graph());
    if (!g) return SHX_ENOMEM;
    shx_system* const sys[1] = {s};
    int rc = shx_batch_create(s->ctx, sys, u0, &t0, &dt, 1, s->march_ty == 8 ? 8 : 4, &g->b);
    if (rc == SHX_OK) rc = shx_batch_capture(g->b, nsteps);
    if (rc != SHX_OK) {
        if (g->b) shx_batch_destroy(g->b);
        return rc;
    }
    *out = g.release();
    return SHX_OK;
}

int shx_rk4_graph_launch(shx_rk4_graph* g) { return g ? shx_batch_launch_async(g->b) : SHX_EINVAL; }

int shx_rk4_graph_time(shx_rk4_graph* g, double* t) { return g && t ? shx_batch_times(g->b, t) : SHX_EINVAL; }

int shx_rk4_graph_destroy(shx_rk4_graph* g)
{
    if (!g) return SHX_EINVAL;
    const int rc = shx_batch_destroy(g->b);
    delete g;
    return rc;
}

}  // extern "C"
