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
// Per rhs: the points without a row of the first direction and the cut
// points of the five rhs fields are zeroed; for each direction, one kernel
// writes the negated fluxes at the cut points (the B terms' values), then one
// pass applies the direction's first derivative to the negated fluxes of the
// five fields at once, computing each tap's flux from the conserved values
// there (interior rows of the fast intervals), and one more evaluates the
// derivative's listed rows (closures with their B terms).  The first
// direction assigns, the others add; a last kernel zeroes the Dirichlet
// faces.  Negation is exact and commutes with rounding to nearest, so summing
// the products of the negated fluxes gives -(the sum of the products): an
// exact zero may take the other sign, no other value changes.  After each
// stage's rhs one pass over all five fields performs the slot operations of
// the RK4 schedule that follow it (rk update, next stage input or result).
// HBM-bound: per direction the five conserved fields are read once (40 B/pt)
// and the five rhs fields written (and read after the first direction).
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
    int* norow = nullptr;  // mesh points with no row of the first direction
    long nnorow = 0;
    double* FR[5][3] = {};  // one direction's negated fluxes at the cut points
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

// The five fluxes along dir at one state, negated, as the rhs takes them:
// the pressure term added to the momentum component of dir after its
// convective part.
__device__ __forceinline__ void neg_flux(double g, int dir, const double (&U)[5], double (&F)[5])
{
    const double r = U[0], mx = U[1], my = U[2], mz = U[3], E = U[4];
    const double pr = (g - 1.0) * (E - 0.5 * (mx * mx + my * my + mz * mz) / r);
    const double md = dir == 0 ? mx : dir == 1 ? my : mz;
    const double ud = md / r;
    double f1 = mx * ud, f2 = my * ud, f3 = mz * ud;
    if (dir == 0) f1 += pr;
    else if (dir == 1) f2 += pr;
    else f3 += pr;
    F[0] = -md;
    F[1] = -f1;
    F[2] = -f2;
    F[3] = -f3;
    F[4] = -((E + pr) * ud);
}

// The negated fluxes along dir at every point of the range of s (field k of
// U is s.a[k], of F s.d[k]).
__global__ void k_flux_neg(Five4 s, int dir, double g)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.off[4]) return;
    const int p = part_of(s, i);
    const long j = i - base_of(s, p);
    double U[5], F[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) U[k] = pick4(p, s.a[k][0], s.a[k][1], s.a[k][2], s.a[k][3])[j];
    neg_flux(g, dir, U, F);
#pragma unroll
    for (int k = 0; k < 5; ++k) pick4(p, s.d[k][0], s.d[k][1], s.d[k][2], s.d[k][3])[j] = F[k];
}

struct FiveC {
    const double* p[5];
};
struct FiveD {
    double* p[5];
};

// rhs_k (op)= D_dir applied to the negated flux of field k, the five fields
// at once, for the interior rows of the fast intervals: each tap's flux is
// computed from the five conserved values there (the same arithmetic as
// k_flux_neg), and the dot products are summed from 0.0 in tap order, as the
// derivative apply sums them over a stored flux.  Rows outside the fast
// intervals come from the derivative's slow-row list (k_euler_dflux_list).
template <int DIR, int W, bool PLUS>
__global__ __launch_bounds__(kBlock) void k_euler_dflux(Tiling t, DevDeriv d, double g, FiveC U, FiveD rhs)
{
    constexpr int hw = W / 2;
    int i, j, k;
    long idx;
    if (!(DIR == 2 ? tile_point_zpatch(t, i, j, k, idx) : tile_point(t, i, j, k, idx))) return;
    const int pair = DIR == 0 ? j * t.n2 + k : DIR == 1 ? i * t.n2 + k : i * t.n1 + j;
    const int q = DIR == 0 ? i : DIR == 1 ? j : k;
    const unsigned long long fd = d.fast[pair];
    double v[W][5];
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const int qq = min(max(q + m - hw, 0), d.nq - 1);
        const long at = idx + (long)(qq - q) * d.stride;
#pragma unroll
        for (int f = 0; f < 5; ++f) v[m][f] = U.p[f][at];
    }
    double old[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) old[f] = PLUS ? rhs.p[f][idx] : 0.0;
    const int lo0 = (int)(fd & 0xffffu), hi0 = (int)((fd >> 16) & 0xffffu);
    const int lo1 = (int)((fd >> 32) & 0xffffu), hi1 = (int)(fd >> 48);
    if (!((q >= lo0 && q < hi0) || (q >= lo1 && q < hi1))) return;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < W; ++m) {
        double F[5];
        neg_flux(g, DIR, v[m], F);
        const double c = d.coef[m];
#pragma unroll
        for (int f = 0; f < 5; ++f) acc[f] += c * F[f];
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) rhs.p[f][idx] = PLUS ? old[f] + acc[f] : acc[f];
}

// The same for rows along x with XP consecutive x planes per lane: the
// negated fluxes of the XP + W - 1 planes the lane's rows read are computed
// once, each from the conserved values of its plane, and shared by the XP
// rows.  t counts groups of XP planes along x.
template <int W, bool PLUS, int XP>
__global__ __launch_bounds__(kBlock) void k_euler_dflux_x(Tiling t, int n0, DevDeriv d, double g, FiveC U,
                                                          FiveD rhs)
{
    constexpr int hw = W / 2;
    int grp, j, k;
    long unused;
    if (!tile_point(t, grp, j, k, unused)) return;
    const long plane = (long)t.n1 * t.n2, col = (long)j * t.n2 + k;
    const int i0 = grp * XP;
    const unsigned long long fd = d.fast[j * t.n2 + k];
    double F[XP + W - 1][5];
#pragma unroll
    for (int m = 0; m < XP + W - 1; ++m) {
        const int qq = min(max(i0 + m - hw, 0), d.nq - 1);
        double Um[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) Um[f] = U.p[f][(long)qq * plane + col];
        neg_flux(g, 0, Um, F[m]);
    }
    const int lo0 = (int)(fd & 0xffffu), hi0 = (int)((fd >> 16) & 0xffffu);
    const int lo1 = (int)((fd >> 32) & 0xffffu), hi1 = (int)(fd >> 48);
#pragma unroll
    for (int p = 0; p < XP; ++p) {
        const int q = i0 + p;
        if (q >= n0) break;
        if (!((q >= lo0 && q < hi0) || (q >= lo1 && q < hi1))) continue;
        const long idx = (long)q * plane + col;
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < W; ++m) {
            const double c = d.coef[m];
#pragma unroll
            for (int f = 0; f < 5; ++f) acc[f] += c * F[p + m][f];
        }
#pragma unroll
        for (int f = 0; f < 5; ++f) rhs.p[f][idx] = PLUS ? rhs.p[f][idx] + acc[f] : acc[f];
    }
}

// The listed rows of the derivative (closures with their B terms, interior
// rows of a line's third and later segments) for the five fields, one lane
// per row: the columns' fluxes computed from U, the B terms from the negated
// fluxes at the cut points of R(dir) (FR), each field summed as
// row_in_block sums it.
template <int W, bool PLUS>
__global__ __launch_bounds__(256) void k_euler_dflux_list(long n, const SlowRow* __restrict__ rows, DevDeriv d,
                                                          int dir, double g, FiveC U, FiveC FR, FiveD rhs)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SlowRow e = rows[i];
    const shx_line_block blk = d.blocks[e.b];
    const long st = d.stride;
    const int lr = e.q - blk.lo;
    const int total = blk.nl + blk.ni + blk.nr;
    const long base = e.idx - (long)e.q * st;
    const double* a;
    long c0;
    int ncols;
    if (lr < blk.nl) {
        a = d.pool + blk.lcoef + (long)lr * blk.lcols;
        c0 = base + (long)blk.lcol0 * st;
        ncols = blk.lcols;
    } else if (lr < blk.nl + blk.ni) {
        a = d.coef;
        c0 = e.idx - (long)(W / 2) * st;
        ncols = W;
    } else {
        const int rr = lr - blk.nl - blk.ni;
        a = d.pool + blk.rcoef + (long)rr * blk.rcols;
        c0 = base + (long)blk.rcol0 * st;
        ncols = blk.rcols;
    }
    double dot[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < ncols; ++c) {
        const long at = c0 + (long)c * st;
        double Uc[5], F[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) Uc[f] = U.p[f][at];
        neg_flux(g, dir, Uc, F);
        const double w = a[c];
#pragma unroll
        for (int f = 0; f < 5; ++f) dot[f] += w * F[f];
    }
    const bool bl = blk.lobj >= 0 && lr < blk.nl;
    const bool br = blk.robj >= 0 && lr >= total - blk.nr;
    const double wl = bl ? d.pool[blk.lbw + lr] : 0.0;
    const double wr = br ? d.pool[blk.rbw + (lr - (total - blk.nr))] : 0.0;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        RowVal r;
        r.dot = dot[f];
        r.b0 = r.b1 = 0.0;
        r.nb = 0;
        r.row = true;
        if (bl) {
            r.b0 = wl * FR.p[f][blk.lobj];
            r.nb = 1;
        }
        if (br) {
            const double v = wr * FR.p[f][blk.robj];
            if (r.nb == 0) r.b0 = v;
            else r.b1 = v;
            r.nb++;
        }
        rhs.p[f][e.idx] = PLUS ? row_plus(rhs.p[f][e.idx], r) : row_eq(r);
    }
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

// d = 0 or d = a on field blockIdx.y of five (slot_ops.hpp:23-59).
enum { kSlotZero, kSlotCopy };
template <int OP>
__global__ void k_slot5(Five4 s, double)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.off[4]) return;
    const int k = (int)blockIdx.y;
    const int p = part_of(s, i);
    const long j = i - base_of(s, p);
    double* d = pick4(p, s.d[k][0], s.d[k][1], s.d[k][2], s.d[k][3]);
    d[j] = OP == kSlotZero ? 0.0 : pick4(p, s.a[k][0], s.a[k][1], s.a[k][2], s.a[k][3])[j];
}

// After stage i's rhs (b = srhs, d = rk, a = u0), in one pass: rk += c1 b
// (rk = 0 + c1 b at the first stage), then the next stage input o = a + c2 b,
// or, after the last stage, o = a + 1.0 rk with the updated rk: the slot
// operations of rk4.cpp:16-47 that follow one another, each with the same
// arithmetic.
enum { kStepFirst, kStepMid, kStepLast };
struct Out4 {
    double* o[5][4];
};
template <int OP>
__global__ void k_step_slots(Five4 s, Out4 o, double c1, double c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.off[4]) return;
    const int k = (int)blockIdx.y;
    const int p = part_of(s, i);
    const long j = i - base_of(s, p);
    double* rk = pick4(p, s.d[k][0], s.d[k][1], s.d[k][2], s.d[k][3]);
    const double u0 = pick4(p, s.a[k][0], s.a[k][1], s.a[k][2], s.a[k][3])[j];
    const double sr = pick4(p, s.b[k][0], s.b[k][1], s.b[k][2], s.b[k][3])[j];
    double* out = pick4(p, o.o[k][0], o.o[k][1], o.o[k][2], o.o[k][3]);
    const double r = (OP == kStepFirst ? 0.0 : rk[j]) + c1 * sr;
    rk[j] = r;
    out[j] = OP == kStepLast ? u0 + 1.0 * r : u0 + c2 * sr;
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

template <int DIR, int W>
void dflux_launch(shx_euler* e, const shx_derivative* d, bool plus, const FiveC& U, const FiveC& FR, const FiveD& R)
{
    hipStream_t st = e->ctx->stream;
    const double g = e->gamma;
    // rows along x: four planes per lane when the grid keeps at least eight
    // workgroups per CU
    const int g4[3] = {(d->n[0] + kDerivPlanes - 1) / kDerivPlanes, d->n[1], d->n[2]};
    const Tiling t4 = make_tiling(g4);
    if (DIR == 0 && t4.nblocks >= 8u * 256u) {
        if (plus) k_euler_dflux_x<W, true, kDerivPlanes><<<t4.nblocks, kBlock, 0, st>>>(t4, d->n[0], d->dev, g, U, R);
        else k_euler_dflux_x<W, false, kDerivPlanes><<<t4.nblocks, kBlock, 0, st>>>(t4, d->n[0], d->dev, g, U, R);
    } else {
        const Tiling t = make_tiling(d->n);
        if (plus) k_euler_dflux<DIR, W, true><<<t.nblocks, kBlock, 0, st>>>(t, d->dev, g, U, R);
        else k_euler_dflux<DIR, W, false><<<t.nblocks, kBlock, 0, st>>>(t, d->dev, g, U, R);
    }
    if (d->nslow == 0) return;
    const unsigned grid = (unsigned)((d->nslow + 255) / 256);
    if (plus) k_euler_dflux_list<W, true><<<grid, 256, 0, st>>>(d->nslow, d->slow, d->dev, DIR, g, U, FR, R);
    else k_euler_dflux_list<W, false><<<grid, 256, 0, st>>>(d->nslow, d->slow, d->dev, DIR, g, U, FR, R);
}

int rhs_launch(shx_euler* e, const shx_field* U, shx_field* rhs)
{
    hipStream_t st = e->ctx->stream;
    // the first direction assigns every mesh point that has a row of it;
    // the others and the cut points start from zero
    {
        Five4 z = five4(e, rhs, nullptr, nullptr);
        for (int k = 0; k < 5; ++k) z.d[k][0] = nullptr;
        z.off[1] = 0;
        for (int c = 0; c < 3; ++c) z.off[2 + c] = z.off[1 + c] + e->nR[c];
        if (z.off[4] > 0) k_slot5<kSlotZero><<<dim3(blocks_for(z.off[4]), 5), kT, 0, st>>>(z, 0.0);
        if (e->nnorow)
            k_zero_at5<<<blocks_for(e->nnorow), kT, 0, st>>>(e->nnorow, e->norow, five4(e, rhs, nullptr, nullptr));
    }
    int rc = check(e->ctx, "rhs fill");
    // negated fluxes at the cut points (the B terms' values): part 0 empty
    Five4 fs = five4(e, nullptr, U, nullptr);
    for (int k = 0; k < 5; ++k) {
        fs.a[k][0] = nullptr;
        for (int c = 0; c < 3; ++c) fs.d[k][1 + c] = e->FR[k][c];
    }
    fs.off[1] = 0;
    for (int c = 0; c < 3; ++c) fs.off[2 + c] = fs.off[1 + c] + e->nR[c];
    FiveC Ud;
    FiveD Rd;
    for (int k = 0; k < 5; ++k) {
        Ud.p[k] = U[k].D;
        Rd.p[k] = rhs[k].D;
    }
    bool first = true;
    for (int dir = 0; dir < 3 && rc == SHX_OK; ++dir) {
        const shx_derivative* d = e->d[dir];
        if (e->n[dir] < 2 || !d) continue;
        if (fs.off[4] > 0) k_flux_neg<<<blocks_for(fs.off[4]), kT, 0, st>>>(fs, dir, e->gamma);
        FiveC FR;
        for (int k = 0; k < 5; ++k) FR.p[k] = e->FR[k][dir];
        const int key = dir * 10 + d->dev.width;
        switch (key) {
        case 3: dflux_launch<0, 3>(e, d, !first, Ud, FR, Rd); break;
        case 5: dflux_launch<0, 5>(e, d, !first, Ud, FR, Rd); break;
        case 13: dflux_launch<1, 3>(e, d, !first, Ud, FR, Rd); break;
        case 15: dflux_launch<1, 5>(e, d, !first, Ud, FR, Rd); break;
        case 23: dflux_launch<2, 3>(e, d, !first, Ud, FR, Rd); break;
        case 25: dflux_launch<2, 5>(e, d, !first, Ud, FR, Rd); break;
        default: rc = fail(e->ctx, SHX_EUNSUPPORTED, "euler: interior width must be 3 or 5");
        }
        if (rc == SHX_OK) rc = check(e->ctx, "euler flux-derivative kernels");
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
    // mesh points without a row of the first present direction (solid
    // points, dropped Dirichlet rows, lines on a Dirichlet face of another
    // direction): the rhs zeroes them, the first direction assigns the rest
    {
        int c0 = -1;
        for (int c = 0; c < 3 && c0 < 0; ++c)
            if (e->n[c] >= 2 && d[c]) c0 = c;
        std::vector<char> row((size_t)e->N, 0);
        if (c0 >= 0) {
            const shx_derivative* dc = d[c0];
            const int nfast = c0 == 2 ? e->n[1] : e->n[2];
            const long st = c0 == 0 ? (long)e->n[1] * e->n[2] : c0 == 1 ? e->n[2] : 1;
            for (const shx_line_block& blk : dc->h_blocks) {
                const int sl = blk.pair / nfast, f = blk.pair % nfast;
                long base;
                if (c0 == 0) base = (long)sl * e->n[2] + f;
                else if (c0 == 1) base = (long)sl * e->n[1] * e->n[2] + f;
                else base = ((long)sl * e->n[1] + f) * e->n[2];
                const int total = blk.nl + blk.ni + blk.nr;
                for (int q = blk.lo; q < blk.lo + total; ++q) row[(size_t)(base + (long)q * st)] = 1;
            }
        }
        std::vector<int> norow;
        for (long i = 0; i < e->N; ++i)
            if (!row[(size_t)i]) norow.push_back((int)i);
        e->nnorow = (long)norow.size();
        if (rc == SHX_OK) rc = upload_n(ctx, &e->norow, norow.data(), norow.size());
    }
    for (int k = 0; k < 5 && rc == SHX_OK; ++k) {
        hipError_t err = hipSuccess;
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
    hipFree(e->norow);
    for (int k = 0; k < 5; ++k) {
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
    // Stage 0 reads u0 itself (the schedule's out = u0 copy is its only
    // use before out is overwritten whole).  After each stage's rhs one pass
    // updates rk (from zero at the first stage) and forms the next stage
    // input, or the step's result after the last stage; srhs is overwritten
    // whole by every rhs.
    int rc = SHX_OK;
    Out4 o;
    for (int k = 0; k < 5; ++k) {
        o.o[k][0] = out[k].D;
        for (int c = 0; c < 3; ++c) o.o[k][1 + c] = out[k].R[c];
    }
    const Five4 sl = five4(e, rk, u0, srhs);
    const dim3 grid(blocks_for(sl.off[4]), 5);
    hipStream_t st = e->ctx->stream;
    for (int i = 0; i < 4 && rc == SHX_OK; ++i) {
        rc = rhs_launch(e, i == 0 ? u0 : out, srhs);
        if (rc != SHX_OK) break;
        const double c1 = dt * rkf[i];
        if (i == 0) k_step_slots<kStepFirst><<<grid, kT, 0, st>>>(sl, o, c1, dt * rki[1]);
        else if (i < 3) k_step_slots<kStepMid><<<grid, kT, 0, st>>>(sl, o, c1, dt * rki[i + 1]);
        else k_step_slots<kStepLast><<<grid, kT, 0, st>>>(sl, o, c1, 1.0);
        rc = check(e->ctx, "euler slot kernels");
        if (rc == SHX_OK) rc = wall_launch(e, out);
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
