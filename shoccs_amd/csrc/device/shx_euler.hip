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
// Per rhs: one fill zeroes the five rhs fields; for each direction, one flux
// kernel over the five fields' mesh and cut points writes the negated fluxes,
// then five derivative applies (O and B in one launch, eq for the first
// direction, plus_eq after) accumulate -D F directly; a last kernel zeroes the
// Dirichlet faces.  Negation is exact and commutes with rounding to nearest,
// so -(sum of products) and the sum of the products of the negated fluxes
// are the same number (an exact zero may take the other sign).  Every slot
// operation of the RK4 schedule is one launch over all five fields.
// HBM-bound: about 5 x (8 reads + 8 writes) B/pt per direction for the
// fluxes and 5 x 24 B/pt per direction for the applies.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "shx_internal.hpp"
#include "shx_state.hpp"

using namespace shx;

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

// The five fields' arrays as one index range [0, off[4]): part 0 holds the
// mesh points [0, off[1]), parts 1-3 the cut points of R(x), R(y), R(z).
// Pointers are picked by comparisons (no dynamic index into a kernel
// argument, which would copy it to scratch memory).
struct Five4 {
    double* d[5][4];
    const double* a[5][4];
    const double* b[5][4];
    long off[5];
};

__device__ __forceinline__ int part_of(const Five4& s, long i)
{
    return (int)(i >= s.off[1]) + (int)(i >= s.off[2]) + (int)(i >= s.off[3]);
}
template <typename T>
__device__ __forceinline__ T pick4(int p, T x0, T x1, T x2, T x3)
{
    return p == 0 ? x0 : p == 1 ? x1 : p == 2 ? x2 : x3;
}
__device__ __forceinline__ long base_of(const Five4& s, int p)
{
    return pick4(p, 0L, s.off[1], s.off[2], s.off[3]);
}

// The five fluxes along dir, negated, at every mesh and cut point: the
// pressure term added to the momentum component of dir after its convective
// part, as the rhs takes them; field k of U is s.a[k], of F s.d[k].
__global__ void k_flux_neg(Five4 s, int dir, double g)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.off[4]) return;
    const int p = part_of(s, i);
    const long j = i - base_of(s, p);
    double U[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) U[k] = pick4(p, s.a[k][0], s.a[k][1], s.a[k][2], s.a[k][3])[j];
    const double r = U[0], mx = U[1], my = U[2], mz = U[3], E = U[4];
    const double pr = (g - 1.0) * (E - 0.5 * (mx * mx + my * my + mz * mz) / r);
    const double md = dir == 0 ? mx : dir == 1 ? my : mz;
    const double ud = md / r;
    double f1 = mx * ud, f2 = my * ud, f3 = mz * ud;
    if (dir == 0) f1 += pr;
    else if (dir == 1) f2 += pr;
    else f3 += pr;
    const double F[5] = {md, f1, f2, f3, (E + pr) * ud};
#pragma unroll
    for (int k = 0; k < 5; ++k) pick4(p, s.d[k][0], s.d[k][1], s.d[k][2], s.d[k][3])[j] = -F[k];
}

// Slip wall: each cut point takes its fluid neighbour's state with the
// normal momentum removed (density and pressure kept).  Cut points of the
// three sets as one range (parts 1-3 of s); field k on D is s.a[k][0].
struct WallArgs {
    const long* fluid[3];
    const double* nrm[3];
};
__global__ void k_wall(Five4 s, WallArgs w, double g1)
{
    const long i = s.off[1] + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.off[4]) return;
    const int p = part_of(s, i);
    const long q = i - base_of(s, p);
    const long f = pick4(p, (const long*)nullptr, w.fluid[0], w.fluid[1], w.fluid[2])[q];
    const double* nr = pick4(p, (const double*)nullptr, w.nrm[0], w.nrm[1], w.nrm[2]) + 3 * q;
    const double n0 = nr[0], n1 = nr[1], n2 = nr[2];
    const double r = s.a[0][0][f], E = s.a[4][0][f];
    double m0 = s.a[1][0][f], m1 = s.a[2][0][f], m2 = s.a[3][0][f];
    const double pr = g1 * (E - 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r);
    const double mn = m0 * n0 + m1 * n1 + m2 * n2;
    m0 = m0 - mn * n0;
    m1 = m1 - mn * n1;
    m2 = m2 - mn * n2;
    const double V[5] = {r, m0, m1, m2, pr / g1 + 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r};
#pragma unroll
    for (int k = 0; k < 5; ++k) pick4(p, s.d[k][0], s.d[k][1], s.d[k][2], s.d[k][3])[q] = V[k];
}

__global__ void k_zero_at5(long n, const int* __restrict__ sel, Five4 s)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long i = sel[q];
#pragma unroll
    for (int k = 0; k < 5; ++k) s.d[k][0][i] = 0.0;
}

// Slot operations of rk4.cpp:16-47 (slot_ops.hpp:23-59) on field
// blockIdx.y of five: d = 0, d = a, d = a + c b, d += c b, d = 0 + c b.
enum { kSlotZero, kSlotCopy, kSlotLc, kSlotAcc, kSlotAccFromZero };
template <int OP>
__global__ void k_slot5(Five4 s, double c)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.off[4]) return;
    const int k = (int)blockIdx.y;
    const int p = part_of(s, i);
    const long j = i - base_of(s, p);
    double* d = pick4(p, s.d[k][0], s.d[k][1], s.d[k][2], s.d[k][3]);
    if (OP == kSlotZero) {
        d[j] = 0.0;
    } else if (OP == kSlotCopy) {
        d[j] = pick4(p, s.a[k][0], s.a[k][1], s.a[k][2], s.a[k][3])[j];
    } else if (OP == kSlotLc) {
        const double x = pick4(p, s.a[k][0], s.a[k][1], s.a[k][2], s.a[k][3])[j];
        d[j] = x + c * pick4(p, s.b[k][0], s.b[k][1], s.b[k][2], s.b[k][3])[j];
    } else if (OP == kSlotAcc) {
        d[j] += c * pick4(p, s.b[k][0], s.b[k][1], s.b[k][2], s.b[k][3])[j];
    } else {
        d[j] = 0.0 + c * pick4(p, s.b[k][0], s.b[k][1], s.b[k][2], s.b[k][3])[j];
    }
}

Five4 five4(const shx_euler* e, shx_field* d, const shx_field* a, const shx_field* b)
{
    Five4 s{};
    for (int k = 0; k < 5; ++k) {
        s.d[k][0] = d ? d[k].D : nullptr;
        s.a[k][0] = a ? a[k].D : nullptr;
        s.b[k][0] = b ? b[k].D : nullptr;
        for (int c = 0; c < 3; ++c) {
            s.d[k][1 + c] = d ? d[k].R[c] : nullptr;
            s.a[k][1 + c] = a ? a[k].R[c] : nullptr;
            s.b[k][1 + c] = b ? b[k].R[c] : nullptr;
        }
    }
    s.off[0] = 0;
    s.off[1] = e->N;
    for (int c = 0; c < 3; ++c) s.off[2 + c] = s.off[1 + c] + e->nR[c];
    return s;
}

template <int OP>
void slot5(shx_euler* e, shx_field* d, const shx_field* a, double c, const shx_field* b)
{
    const Five4 s = five4(e, d, a, b);
    k_slot5<OP><<<dim3(blocks_for(s.off[4]), 5), kT, 0, e->ctx->stream>>>(s, c);
}

int wall_launch(shx_euler* e, shx_field* U)
{
    const Five4 s = five4(e, U, U, nullptr);
    const long n = s.off[4] - s.off[1];
    if (n == 0) return SHX_OK;
    WallArgs w;
    for (int c = 0; c < 3; ++c) {
        w.fluid[c] = e->wall_fluid[c];
        w.nrm[c] = e->wall_normal[c];
    }
    k_wall<<<blocks_for(n), kT, 0, e->ctx->stream>>>(s, w, e->gamma - 1.0);
    return check(e->ctx, "wall kernel");
}

int rhs_launch(shx_euler* e, const shx_field* U, shx_field* rhs)
{
    hipStream_t st = e->ctx->stream;
    slot5<kSlotZero>(e, rhs, nullptr, 0.0, nullptr);
    int rc = check(e->ctx, "rhs fill");
    bool first = true;
    Five4 fs = five4(e, nullptr, U, nullptr);
    for (int k = 0; k < 5; ++k) {
        fs.d[k][0] = e->F[k];
        for (int c = 0; c < 3; ++c) fs.d[k][1 + c] = e->FR[k][c];
    }
    for (int dir = 0; dir < 3 && rc == SHX_OK; ++dir) {
        if (e->n[dir] < 2 || !e->d[dir]) continue;
        k_flux_neg<<<blocks_for(fs.off[4]), kT, 0, st>>>(fs, dir, e->gamma);
        rc = check(e->ctx, "flux kernel");
        for (int k = 0; k < 5 && rc == SHX_OK; ++k)
            rc = launch_derivative(e->d[dir], e->F[k], e->FR[k][dir], rhs[k].D, !first, true);
        first = false;
    }
    if (rc != SHX_OK) return rc;
    if (e->nfaces) k_zero_at5<<<blocks_for(e->nfaces), kT, 0, st>>>(e->nfaces, e->faces, five4(e, rhs, nullptr, nullptr));
    return check(e->ctx, "rhs finish kernels");
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
    // rk starts from zero (its first update is 0 + c srhs); srhs is
    // overwritten whole by every rhs
    slot5<kSlotCopy>(e, out, u0, 0.0, nullptr);
    int rc = check(e->ctx, "euler copy");
    for (int i = 0; i < 4 && rc == SHX_OK; ++i) {
        if (i > 0) {
            slot5<kSlotLc>(e, out, u0, dt * rki[i], srhs);
            rc = wall_launch(e, out);
        }
        if (rc == SHX_OK) rc = rhs_launch(e, out, srhs);
        if (rc == SHX_OK) {
            if (i == 0) slot5<kSlotAccFromZero>(e, rk, nullptr, dt * rkf[i], srhs);
            else slot5<kSlotAcc>(e, rk, nullptr, dt * rkf[i], srhs);
            rc = check(e->ctx, "euler slot kernels");
        }
    }
    if (rc == SHX_OK) {
        slot5<kSlotLc>(e, out, u0, 1.0, rk);
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
        if (rc == SHX_OK) {
            slot5<kSlotCopy>(e, u0, out, 0.0, nullptr);
            rc = check(e->ctx, "euler copy");
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
