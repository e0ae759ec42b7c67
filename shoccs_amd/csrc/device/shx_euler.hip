// shx_euler.hip -- the compressible Euler system on one GPU: five conserved
// fields (rho, rho u, rho v, rho w, E), rhs_k = -(D_x F_x,k + D_y F_y,k +
// D_z F_z,k) with the first-derivative operators of the problem, a slip wall
// at the bodies' cut points, Dirichlet grid faces held, and the RK4 slot
// schedule of src/temporal/rk4.cpp:16-47 over the five fields.
//
// The reference has no Euler system (src/systems/inviscid_vortex.cpp:118-154
// is a stub); the discretisation is this library's own, assembled from the
// reference's operators, and every step is plain FP64 arithmetic (no
// transcendentals), so a CPU restatement of it agrees to the last bit.
//
// Per rhs: for each direction, one flux kernel (five fields, mesh and cut
// points), then five derivative applies (O and B in one launch, eq for the
// first direction, plus_eq after); then one kernel negates, zeroes the
// Dirichlet faces and the cut points.  HBM-bound: about 5 x (8 reads + 8
// writes) B/pt per direction for the fluxes and 5 x 24 B/pt per direction for
// the applies.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "shx_internal.hpp"
#include "shx_state.hpp"

using namespace shx;

struct FivePtr {
    double* p[5];
};
struct FiveCPtr {
    const double* p[5];
};

struct shx_euler {
    shx_ctx* ctx = nullptr;
    shx_derivative* d[3] = {nullptr, nullptr, nullptr};
    int n[3] = {1, 1, 1};
    long N = 0;
    long nR[3] = {0, 0, 0};
    double gamma = 1.4;
    long* wall_fluid[3] = {nullptr, nullptr, nullptr};
    double* wall_normal[3] = {nullptr, nullptr, nullptr};
    int* faces = nullptr;  // Dirichlet grid-face points
    long nfaces = 0;
    double* F[5] = {};     // one direction's fluxes on D
    double* FR[5][3] = {};  // and on the cut points
};

namespace {

constexpr int kT = 256;
inline unsigned blocks_for(long n) { return (unsigned)((n + kT - 1) / kT); }

int check(shx_ctx* ctx, const char* what)
{
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, what);
}

// The five fluxes along dir at n points (mesh or cut points), the pressure
// term added to the momentum component of dir after its convective part.
__global__ void k_flux(long n, int dir, double g, FiveCPtr U, FivePtr F)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = U.p[0][i], mx = U.p[1][i], my = U.p[2][i], mz = U.p[3][i], E = U.p[4][i];
    const double p = (g - 1.0) * (E - 0.5 * (mx * mx + my * my + mz * mz) / r);
    const double md = dir == 0 ? mx : dir == 1 ? my : mz;
    const double ud = md / r;
    double f1 = mx * ud, f2 = my * ud, f3 = mz * ud;
    if (dir == 0) f1 += p;
    else if (dir == 1) f2 += p;
    else f3 += p;
    F.p[0][i] = md;
    F.p[1][i] = f1;
    F.p[2][i] = f2;
    F.p[3][i] = f3;
    F.p[4][i] = (E + p) * ud;
}

// Slip wall: each cut point takes its fluid neighbour's state with the
// normal momentum removed (density and pressure kept).
__global__ void k_wall(long n, double g1, const long* __restrict__ fluid, const double* __restrict__ nrm,
                       FiveCPtr UD, FivePtr UR)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long f = fluid[i];
    const double n0 = nrm[3 * i], n1 = nrm[3 * i + 1], n2 = nrm[3 * i + 2];
    const double r = UD.p[0][f], E = UD.p[4][f];
    double m0 = UD.p[1][f], m1 = UD.p[2][f], m2 = UD.p[3][f];
    const double p = g1 * (E - 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r);
    const double mn = m0 * n0 + m1 * n1 + m2 * n2;
    m0 = m0 - mn * n0;
    m1 = m1 - mn * n1;
    m2 = m2 - mn * n2;
    UR.p[0][i] = r;
    UR.p[1][i] = m0;
    UR.p[2][i] = m1;
    UR.p[3][i] = m2;
    UR.p[4][i] = p / g1 + 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r;
}

__global__ void k_negate5(long n, FivePtr x)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 5; ++k) x.p[k][i] = -x.p[k][i];
}

__global__ void k_zero_at5(long n, const int* __restrict__ sel, FivePtr x)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long i = sel[q];
#pragma unroll
    for (int k = 0; k < 5; ++k) x.p[k][i] = 0.0;
}

// slot_ops.hpp:23-59
__global__ void k_lc(long n, double* __restrict__ d, const double* __restrict__ s0, double c,
                     const double* __restrict__ r)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s0[i] + c * r[i];
}

__global__ void k_acc(long n, double* __restrict__ d, double c, const double* __restrict__ r)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] += c * r[i];
}

FiveCPtr cD(const shx_field* U)
{
    FiveCPtr p;
    for (int k = 0; k < 5; ++k) p.p[k] = U[k].D;
    return p;
}
FivePtr mD(shx_field* U)
{
    FivePtr p;
    for (int k = 0; k < 5; ++k) p.p[k] = U[k].D;
    return p;
}

int zero5(shx_euler* e, shx_field* X)
{
    int rc = SHX_OK;
    for (int k = 0; k < 5 && rc == SHX_OK; ++k) {
        rc = launch_fill(e->ctx, X[k].D, 0.0, e->N);
        for (int c = 0; c < 3 && rc == SHX_OK; ++c)
            if (e->nR[c] > 0) rc = launch_fill(e->ctx, X[k].R[c], 0.0, e->nR[c]);
    }
    return rc;
}

int wall_launch(shx_euler* e, shx_field* U)
{
    hipStream_t st = e->ctx->stream;
    for (int c = 0; c < 3; ++c) {
        if (e->nR[c] == 0) continue;
        FivePtr R;
        for (int k = 0; k < 5; ++k) R.p[k] = U[k].R[c];
        k_wall<<<blocks_for(e->nR[c]), kT, 0, st>>>(e->nR[c], e->gamma - 1.0, e->wall_fluid[c], e->wall_normal[c],
                                                     cD(U), R);
    }
    return check(e->ctx, "wall kernel");
}

int rhs_launch(shx_euler* e, const shx_field* U, shx_field* rhs)
{
    hipStream_t st = e->ctx->stream;
    int rc = zero5(e, rhs);
    bool first = true;
    for (int dir = 0; dir < 3 && rc == SHX_OK; ++dir) {
        if (e->n[dir] < 2 || !e->d[dir]) continue;
        FivePtr F;
        for (int k = 0; k < 5; ++k) F.p[k] = e->F[k];
        k_flux<<<blocks_for(e->N), kT, 0, st>>>(e->N, dir, e->gamma, cD(U), F);
        for (int c = 0; c < 3; ++c) {
            if (e->nR[c] == 0) continue;
            FiveCPtr UR;
            FivePtr FR;
            for (int k = 0; k < 5; ++k) {
                UR.p[k] = U[k].R[c];
                FR.p[k] = e->FR[k][c];
            }
            k_flux<<<blocks_for(e->nR[c]), kT, 0, st>>>(e->nR[c], dir, e->gamma, UR, FR);
        }
        rc = check(e->ctx, "flux kernel");
        for (int k = 0; k < 5 && rc == SHX_OK; ++k)
            rc = launch_derivative(e->d[dir], e->F[k], e->FR[k][dir], rhs[k].D, !first, true);
        first = false;
    }
    if (rc != SHX_OK) return rc;
    k_negate5<<<blocks_for(e->N), kT, 0, st>>>(e->N, mD(rhs));
    if (e->nfaces) k_zero_at5<<<blocks_for(e->nfaces), kT, 0, st>>>(e->nfaces, e->faces, mD(rhs));
    return check(e->ctx, "rhs finish kernels");
}

// d = s0 + c r (every component, mesh and cut points)
void lc5(shx_euler* e, shx_field* d, const shx_field* s0, double c, const shx_field* r)
{
    hipStream_t st = e->ctx->stream;
    for (int k = 0; k < 5; ++k) {
        k_lc<<<blocks_for(e->N), kT, 0, st>>>(e->N, d[k].D, s0[k].D, c, r[k].D);
        for (int q = 0; q < 3; ++q)
            if (e->nR[q] > 0) k_lc<<<blocks_for(e->nR[q]), kT, 0, st>>>(e->nR[q], d[k].R[q], s0[k].R[q], c, r[k].R[q]);
    }
}

void acc5(shx_euler* e, shx_field* d, double c, const shx_field* r)
{
    hipStream_t st = e->ctx->stream;
    for (int k = 0; k < 5; ++k) {
        k_acc<<<blocks_for(e->N), kT, 0, st>>>(e->N, d[k].D, c, r[k].D);
        for (int q = 0; q < 3; ++q)
            if (e->nR[q] > 0) k_acc<<<blocks_for(e->nR[q]), kT, 0, st>>>(e->nR[q], d[k].R[q], c, r[k].R[q]);
    }
}

template <typename T>
int upload_n(shx_ctx* ctx, T** dst, const T* src, size_t n)
{
    *dst = nullptr;
    if (n == 0) return SHX_OK;
    hipError_t err = hipMalloc((void**)dst, n * sizeof(T));
    if (err == hipSuccess) err = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
    return err == hipSuccess ? SHX_OK : hip_fail(ctx, err, "euler upload");
}

}  // namespace

extern "C" {

int shx_euler_create(shx_ctx* ctx, shx_derivative* const d[3], const shx_euler_desc* desc, shx_euler** out)
{
    if (!ctx || !d || !desc || !out) return SHX_EINVAL;
    *out = nullptr;
    shx_derivative* any = d[0] ? d[0] : d[1] ? d[1] : d[2];
    if (!any) return SHX_EINVAL;
    if (!(desc->gamma > 1.0)) return fail(ctx, SHX_EINVAL, "euler: gamma must exceed 1");
    for (int c = 0; c < 3; ++c)
        if (d[c] && d[c]->nrrow > 0)
            return fail(ctx, SHX_EUNSUPPORTED, "euler: bodies are slip walls (Dirichlet in the operators' sense)");
    auto* e = new (std::nothrow) shx_euler();
    if (!e) return SHX_ENOMEM;
    e->ctx = ctx;
    e->gamma = desc->gamma;
    for (int c = 0; c < 3; ++c) {
        e->d[c] = d[c];
        e->n[c] = any->n[c];
        e->nR[c] = any->nR[c];
    }
    e->N = (long)e->n[0] * e->n[1] * e->n[2];
    int rc = SHX_OK;
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
        if (e->nR[c] == 0) continue;
        for (long i = 0; i < e->nR[c] && rc == SHX_OK; ++i)
            if (desc->wall_fluid[c][i] < 0 || desc->wall_fluid[c][i] >= e->N)
                rc = fail(ctx, SHX_EINVAL, "euler: wall neighbour outside the mesh");
        if (rc == SHX_OK) rc = upload_n(ctx, &e->wall_fluid[c], (const long*)desc->wall_fluid[c], (size_t)e->nR[c]);
        if (rc == SHX_OK) rc = upload_n(ctx, &e->wall_normal[c], desc->wall_normal[c], (size_t)(3 * e->nR[c]));
    }
    // Dirichlet grid faces hold their state
    std::vector<int> faces;
    {
        const int nx = e->n[0], ny = e->n[1], nz = e->n[2];
        std::vector<char> on((size_t)e->N, 0);
        for (int a = 0; a < 3; ++a)
            for (int side = 0; side < 2; ++side) {
                if (desc->grid_bc[2 * a + side] != SHX_BC_DIRICHLET) continue;
                const int p = side ? e->n[a] - 1 : 0;
                for (int i = 0; i < nx; ++i)
                    for (int j = 0; j < ny; ++j)
                        for (int k = 0; k < nz; ++k) {
                            const int q[3] = {i, j, k};
                            if (q[a] == p) on[((size_t)i * ny + j) * nz + k] = 1;
                        }
            }
        for (long i = 0; i < e->N; ++i)
            if (on[(size_t)i]) faces.push_back((int)i);
    }
    e->nfaces = (long)faces.size();
    if (rc == SHX_OK) rc = upload_n(ctx, &e->faces, faces.data(), faces.size());
    for (int k = 0; k < 5 && rc == SHX_OK; ++k) {
        hipError_t err = hipMalloc((void**)&e->F[k], sizeof(double) * (size_t)e->N);
        for (int c = 0; c < 3 && err == hipSuccess; ++c)
            if (e->nR[c] > 0) err = hipMalloc((void**)&e->FR[k][c], sizeof(double) * (size_t)e->nR[c]);
        if (err != hipSuccess) rc = hip_fail(ctx, err, "euler flux buffers");
    }
    if (rc != SHX_OK) {
        shx_euler_destroy(e);
        return rc;
    }
    *out = e;
    return SHX_OK;
}

int shx_euler_destroy(shx_euler* e)
{
    if (!e) return SHX_EINVAL;
    hipStreamSynchronize(e->ctx->stream);
    for (int c = 0; c < 3; ++c) {
        hipFree(e->wall_fluid[c]);
        hipFree(e->wall_normal[c]);
    }
    hipFree(e->faces);
    for (int k = 0; k < 5; ++k) {
        hipFree(e->F[k]);
        for (int c = 0; c < 3; ++c) hipFree(e->FR[k][c]);
    }
    delete e;
    return SHX_OK;
}

int shx_euler_wall(shx_euler* e, shx_field* U)
{
    if (!e || !U) return SHX_EINVAL;
    const int rc = wall_launch(e, U);
    return rc == SHX_OK ? sync(e->ctx) : rc;
}

int shx_euler_rhs(shx_euler* e, const shx_field* U, shx_field* rhs)
{
    if (!e || !U || !rhs) return SHX_EINVAL;
    const int rc = rhs_launch(e, U, rhs);
    return rc == SHX_OK ? sync(e->ctx) : rc;
}

int shx_euler_rk4_step(shx_euler* e, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs, double dt)
{
    if (!e || !u0 || !out || !rk || !srhs) return SHX_EINVAL;
    static const double rki[4] = {0.0, 0.5, 0.5, 1.0};
    static const double rkf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    int rc = zero5(e, rk);
    if (rc == SHX_OK) rc = zero5(e, srhs);
    for (int k = 0; k < 5 && rc == SHX_OK; ++k) {
        hipMemcpyAsync(out[k].D, u0[k].D, sizeof(double) * (size_t)e->N, hipMemcpyDeviceToDevice, e->ctx->stream);
        for (int q = 0; q < 3; ++q)
            if (e->nR[q] > 0)
                hipMemcpyAsync(out[k].R[q], u0[k].R[q], sizeof(double) * (size_t)e->nR[q], hipMemcpyDeviceToDevice,
                               e->ctx->stream);
        rc = check(e->ctx, "euler copy");
    }
    for (int i = 0; i < 4 && rc == SHX_OK; ++i) {
        if (i > 0) {
            lc5(e, out, u0, dt * rki[i], srhs);
            rc = wall_launch(e, out);
        }
        if (rc == SHX_OK) rc = rhs_launch(e, out, srhs);
        if (rc == SHX_OK) {
            acc5(e, rk, dt * rkf[i], srhs);
            rc = check(e->ctx, "euler slot kernels");
        }
    }
    if (rc == SHX_OK) {
        lc5(e, out, u0, 1.0, rk);
        rc = wall_launch(e, out);
    }
    return rc == SHX_OK ? sync(e->ctx) : rc;
}

int shx_time_euler_steps(shx_euler* e, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs, double dt,
                         int nsteps, double* ms_per_step)
{
    if (!e || nsteps < 1 || !ms_per_step) return SHX_EINVAL;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return hip_fail(e->ctx, hipGetLastError(), "event");
    if (hipEventCreate(&b) != hipSuccess) {
        hipEventDestroy(a);
        return hip_fail(e->ctx, hipGetLastError(), "event");
    }
    int rc = sync(e->ctx);
    hipEventRecord(a, e->ctx->stream);
    for (int s = 0; s < nsteps && rc == SHX_OK; ++s) {
        rc = shx_euler_rk4_step(e, u0, out, rk, srhs, dt);
        // u0 <- out for the next step (simulation_cycle.cpp:61)
        for (int k = 0; k < 5 && rc == SHX_OK; ++k) {
            hipMemcpyAsync(u0[k].D, out[k].D, sizeof(double) * (size_t)e->N, hipMemcpyDeviceToDevice, e->ctx->stream);
            for (int q = 0; q < 3; ++q)
                if (e->nR[q] > 0)
                    hipMemcpyAsync(u0[k].R[q], out[k].R[q], sizeof(double) * (size_t)e->nR[q],
                                   hipMemcpyDeviceToDevice, e->ctx->stream);
        }
    }
    hipEventRecord(b, e->ctx->stream);
    hipError_t err = hipEventSynchronize(b);
    float ms = 0.f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc != SHX_OK) return rc;
    if (err != hipSuccess) return hip_fail(e->ctx, err, "time euler");
    *ms_per_step = (double)ms / nsteps;
    return SHX_OK;
}

}  // extern "C"
