// shx_euler.hip -- the compressible Euler system on one GPU: five conserved
// fields (rho, rho u, rho v, rho w, E), rhs_k = -(D_x F_x,k + D_y F_y,k +
// D_z F_z,k) with the first-derivative operators of the problem, a slip wall
// at the bodies' cut points, Dirichlet grid faces held, and the RK4 slot
// schedule of src/temporal/rk4.cpp:16-47 over the five fields.
//
// The reference has no Euler system (src/systems/inviscid_vortex.cpp:118-154
// is a stub); the discretisation is this library's own, assembled from the
// reference's operators.  In the FP64 form every step is plain FP64
// arithmetic (no transcendentals), so a CPU restatement of it agrees to the
// last bit.
//
// One rhs: the points without a row of the first direction are zeroed; for
// each direction one kernel writes the negated fluxes at the cut points (the
// B terms' values), one pass applies the direction's first derivative to the
// negated fluxes of the five fields at once, computing each tap's flux from
// the conserved values there (interior rows of the fast intervals), and one
// more evaluates the derivative's listed rows (closures with their B terms).
// The first direction assigns, the others add.  Negation is exact and
// commutes with rounding to nearest, so summing the products of the negated
// fluxes gives -(the sum of the products): an exact zero may take the other
// sign, no other value changes.
//
// One RK4 step: per stage the rhs as above, except that the last direction's
// interior pass also finishes the stage: it adds its rows to the partial rhs
// of the point (its listed rows are added before it), zeroes the Dirichlet
// faces, and performs the slot operations of the schedule that follow the
// rhs (rk update, next stage input or result).  The stage input is read at
// the neighbours of a point while the next one is written, so the two live
// in different buffers: the next stage input overwrites the partial rhs of
// its own point.  Then the wall sets the cut points from the new input.
//
// The mixed form keeps the stage inputs after the first and the partial rhs
// in FP32 at the mesh points; the interior flux-derivative rows are computed
// and summed in FP32 with Kahan compensation.  Closure rows, B terms, the cut
// points, u0 and the RK accumulator stay FP64.
#include <hip/hip_runtime.h>

#include <new>
#include <type_traits>
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
    int precision = SHX_F64;
    unsigned face_mask = 0;  // bit 2a + side: that grid face is Dirichlet
    long* wall_fluid[3] = {nullptr, nullptr, nullptr};
    double* wall_normal[3] = {nullptr, nullptr, nullptr};
    int* faces = nullptr;  // Dirichlet grid-face points
    long nfaces = 0;
    int* norow = nullptr;  // mesh points with no row of the first direction
    long nnorow = 0;
    double* FR[5][3] = {};  // one direction's negated fluxes at the cut points
    float* mixF[2][5] = {};  // mixed form: the two FP32 stage/partial-rhs buffers (mesh points)
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

// Five arrays of one type.
template <typename T>
struct P5 {
    T* p[5];
};

// Number types of a pass: TU the stage input at the mesh points, TR the
// partial rhs, TC the arithmetic of the interior rows.
template <typename TU_, typename TR_, typename TC_>
struct Form {
    using TU = TU_;
    using TR = TR_;
    using TC = TC_;
};
using F64 = Form<double, double, double>;
using Mixed0 = Form<double, float, float>;  // mixed form, first stage (its input is u0)
using Mixed = Form<float, float, float>;

// The five fluxes along dir at one state, negated, as the rhs takes them:
// the pressure term added to the momentum component of dir after its
// convective part.
template <typename TC>
__device__ __forceinline__ void neg_flux(TC g, int dir, const TC (&U)[5], TC (&F)[5])
{
    const TC r = U[0], mx = U[1], my = U[2], mz = U[3], E = U[4];
    const TC pr = (g - TC(1)) * (E - TC(0.5) * (mx * mx + my * my + mz * mz) / r);
    const TC md = dir == 0 ? mx : dir == 1 ? my : mz;
    const TC ud = md / r;
    TC f1 = mx * ud, f2 = my * ud, f3 = mz * ud;
    if (dir == 0) f1 += pr;
    else if (dir == 1) f2 += pr;
    else f3 += pr;
    F[0] = -md;
    F[1] = -f1;
    F[2] = -f2;
    F[3] = -f3;
    F[4] = -((E + pr) * ud);
}

// An interior row's sum of products: FP64 products summed from 0.0 in tap
// order (the derivative apply's order); FP32 products summed in FP32 with
// Kahan compensation.
template <typename TC>
struct RowSum {
    TC s = TC(0), c = TC(0);
    __device__ __forceinline__ void add(TC x)
    {
        if constexpr (std::is_same_v<TC, double>) {
            s += x;
        } else {
            const TC y = x - c;
            const TC t = s + y;
            c = (t - s) - y;
            s = t;
        }
    }
};

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

__device__ __forceinline__ bool in_fast(unsigned long long fd, int q)
{
    const int lo0 = (int)(fd & 0xffffu), hi0 = (int)((fd >> 16) & 0xffffu);
    const int lo1 = (int)((fd >> 32) & 0xffffu), hi1 = (int)(fd >> 48);
    return (q >= lo0 && q < hi0) || (q >= lo1 && q < hi1);
}

// rhs_k (op)= D_dir applied to the negated flux of field k, the five fields
// at once, for the interior rows of the fast intervals: each tap's flux is
// computed from the five conserved values there (the same arithmetic as
// k_flux_neg), and the products are summed as RowSum sums them.  Rows outside
// the fast intervals come from the derivative's listed rows
// (k_euler_dflux_list).
template <int DIR, int W, bool PLUS, typename FM>
__global__ __launch_bounds__(kBlock) void k_euler_dflux(Tiling t, DevDeriv d, double g, P5<const typename FM::TU> U,
                                                        P5<typename FM::TR> rhs)
{
    using TC = typename FM::TC;
    using TR = typename FM::TR;
    constexpr int hw = W / 2;
    int i, j, k;
    long idx;
    if (!(DIR == 2 ? tile_point_zpatch(t, i, j, k, idx) : tile_point(t, i, j, k, idx))) return;
    const int pair = DIR == 0 ? j * t.n2 + k : DIR == 1 ? i * t.n2 + k : i * t.n1 + j;
    const int q = DIR == 0 ? i : DIR == 1 ? j : k;
    const unsigned long long fd = d.fast[pair];
    TC v[W][5];
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const int qq = min(max(q + m - hw, 0), d.nq - 1);
        const long at = idx + (long)(qq - q) * d.stride;
#pragma unroll
        for (int f = 0; f < 5; ++f) v[m][f] = (TC)U.p[f][at];
    }
    TR old[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) old[f] = PLUS ? rhs.p[f][idx] : TR(0);
    if (!in_fast(fd, q)) return;
    RowSum<TC> acc[5];
#pragma unroll
    for (int m = 0; m < W; ++m) {
        TC F[5];
        neg_flux((TC)g, DIR, v[m], F);
        const TC c = (TC)d.coef[m];
#pragma unroll
        for (int f = 0; f < 5; ++f) acc[f].add(c * F[f]);
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        if constexpr (std::is_same_v<TR, double>) rhs.p[f][idx] = PLUS ? old[f] + acc[f].s : acc[f].s;
        else rhs.p[f][idx] = PLUS ? (TR)((double)old[f] + (double)acc[f].s) : (TR)acc[f].s;
    }
}

// The same for rows along x with XP consecutive x planes per lane: the
// negated fluxes of the XP + W - 1 planes the lane's rows read are computed
// once, each from the conserved values of its plane, and shared by the XP
// rows.  t counts groups of XP planes along x.
template <int W, bool PLUS, int XP, typename FM>
__global__ __launch_bounds__(kBlock) void k_euler_dflux_x(Tiling t, int n0, DevDeriv d, double g,
                                                          P5<const typename FM::TU> U, P5<typename FM::TR> rhs)
{
    using TC = typename FM::TC;
    using TR = typename FM::TR;
    constexpr int hw = W / 2;
    int grp, j, k;
    long unused;
    if (!tile_point(t, grp, j, k, unused)) return;
    const long plane = (long)t.n1 * t.n2, col = (long)j * t.n2 + k;
    const int i0 = grp * XP;
    const unsigned long long fd = d.fast[j * t.n2 + k];
    TC F[XP + W - 1][5];
#pragma unroll
    for (int m = 0; m < XP + W - 1; ++m) {
        const int qq = min(max(i0 + m - hw, 0), d.nq - 1);
        TC Um[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) Um[f] = (TC)U.p[f][(long)qq * plane + col];
        neg_flux((TC)g, 0, Um, F[m]);
    }
#pragma unroll
    for (int p = 0; p < XP; ++p) {
        const int q = i0 + p;
        if (q >= n0) break;
        if (!in_fast(fd, q)) continue;
        const long idx = (long)q * plane + col;
        RowSum<TC> acc[5];
#pragma unroll
        for (int m = 0; m < W; ++m) {
            const TC c = (TC)d.coef[m];
#pragma unroll
            for (int f = 0; f < 5; ++f) acc[f].add(c * F[p + m][f]);
        }
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            if constexpr (std::is_same_v<TR, double>) rhs.p[f][idx] = PLUS ? rhs.p[f][idx] + acc[f].s : acc[f].s;
            else rhs.p[f][idx] = PLUS ? (TR)((double)rhs.p[f][idx] + (double)acc[f].s) : (TR)acc[f].s;
        }
    }
}

// The listed rows of the derivative (closures with their B terms, interior
// rows of a line's third and later segments) for the five fields, one lane
// per row, in FP64 in both forms: the columns' fluxes computed from U, the B
// terms from the negated fluxes at the cut points of R(dir) (FR), each field
// summed as row_in_block sums it.
template <int W, bool PLUS, typename FM>
__global__ __launch_bounds__(256) void k_euler_dflux_list(long n, const SlowRow* __restrict__ rows, DevDeriv d,
                                                          int dir, double g, P5<const typename FM::TU> U,
                                                          P5<const double> FR, P5<typename FM::TR> rhs)
{
    using TR = typename FM::TR;
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
        for (int f = 0; f < 5; ++f) Uc[f] = (double)U.p[f][at];
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
        rhs.p[f][e.idx] = (TR)(PLUS ? row_plus((double)rhs.p[f][e.idx], r) : row_eq(r));
    }
}

// After stage i's rhs (rk4.cpp:16-47): rk += c1 rhs (rk = 0 + c1 rhs at the
// first stage), then the next stage input o = u0 + c2 rhs, or, after the last
// stage, the result o = u0 + 1.0 rk with the updated rk (which the last
// stage does not store: rk is the schedule's scratch register).
enum { kStepFirst, kStepMid, kStepLast };

template <typename TO>
struct StageOut {
    const double* u0[5];  // unused at the first stage: the input is u0
    double* rk[5];
    TO* out[5];
    double c1, c2;
};

// The last direction's pass of a stage: rhs_k = partial_k + (the interior
// row of D_dir at the point, when it has one), zero on the Dirichlet grid
// faces (face_mask bit 2a + side), then the slot operations.  `rhs` holds the
// partial sums of the earlier directions and this direction's listed rows;
// `so.out` may be `rhs` itself (read, then written, at the lane's own point)
// but never the stage input U, which the taps read at neighbouring points.
template <int DIR, int W, int OP, typename FM, typename TO>
__global__ __launch_bounds__(kBlock) void k_euler_dflux_stage(Tiling t, DevDeriv d, double g,
                                                              P5<const typename FM::TU> U, P5<typename FM::TR> rhs,
                                                              StageOut<TO> so, unsigned face_mask)
{
    using TC = typename FM::TC;
    constexpr int hw = W / 2;
    static_assert(OP != kStepFirst || std::is_same_v<typename FM::TU, double>,
                  "the first stage reads u0 (FP64) as its input");
    int i, j, k;
    long idx;
    if (!(DIR == 2 ? tile_point_zpatch(t, i, j, k, idx) : tile_point(t, i, j, k, idx))) return;
    const int pair = DIR == 0 ? j * t.n2 + k : DIR == 1 ? i * t.n2 + k : i * t.n1 + j;
    const int q = DIR == 0 ? i : DIR == 1 ? j : k;
    const unsigned long long fd = d.fast[pair];
    TC v[W][5];
    double centre[5];
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const int qq = min(max(q + m - hw, 0), d.nq - 1);
        const long at = idx + (long)(qq - q) * d.stride;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const auto x = U.p[f][at];
            v[m][f] = (TC)x;
            if (m == hw) centre[f] = (double)x;
        }
    }
    double r[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) r[f] = (double)rhs.p[f][idx];
    if (in_fast(fd, q)) {
        RowSum<TC> acc[5];
#pragma unroll
        for (int m = 0; m < W; ++m) {
            TC F[5];
            neg_flux((TC)g, DIR, v[m], F);
            const TC c = (TC)d.coef[m];
#pragma unroll
            for (int f = 0; f < 5; ++f) acc[f].add(c * F[f]);
        }
#pragma unroll
        for (int f = 0; f < 5; ++f) r[f] = r[f] + (double)acc[f].s;
    }
    const bool face = ((face_mask & 1u) && i == 0) || ((face_mask & 2u) && i == t.n0 - 1) ||
                      ((face_mask & 4u) && j == 0) || ((face_mask & 8u) && j == t.n1 - 1) ||
                      ((face_mask & 16u) && k == 0) || ((face_mask & 32u) && k == t.n2 - 1);
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        const double rf = face ? 0.0 : r[f];
        const double u0 = OP == kStepFirst ? centre[f] : so.u0[f][idx];
        const double rk = (OP == kStepFirst ? 0.0 : so.rk[f][idx]) + so.c1 * rf;
        if (OP != kStepLast) so.rk[f][idx] = rk;
        so.out[f][idx] = (TO)(OP == kStepLast ? u0 + 1.0 * rk : u0 + so.c2 * rf);
    }
}

// Slip wall: each cut point takes its fluid neighbour's state with the
// normal momentum removed (density and pressure kept), in FP64.  The cut
// points of the three sets form one range: [off[c], off[c + 1]) is R(c).
struct WallArgs {
    const long* fluid[3];
    const double* nrm[3];
};
template <typename TD>
struct WallIO {
    const TD* D[5];
    double* R[5][3];
    long off[4];
};
template <typename TD>
__global__ void k_wall(WallIO<TD> io, WallArgs w, double g1)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= io.off[3]) return;
    const int c = (int)(i >= io.off[1]) + (int)(i >= io.off[2]);
    const long q = i - (c == 0 ? 0L : c == 1 ? io.off[1] : io.off[2]);
    const long f = (c == 0 ? w.fluid[0] : c == 1 ? w.fluid[1] : w.fluid[2])[q];
    const double* nr = (c == 0 ? w.nrm[0] : c == 1 ? w.nrm[1] : w.nrm[2]) + 3 * q;
    const double n0 = nr[0], n1 = nr[1], n2 = nr[2];
    const double r = (double)io.D[0][f], E = (double)io.D[4][f];
    double m0 = (double)io.D[1][f], m1 = (double)io.D[2][f], m2 = (double)io.D[3][f];
    const double pr = g1 * (E - 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r);
    const double mn = m0 * n0 + m1 * n1 + m2 * n2;
    m0 = m0 - mn * n0;
    m1 = m1 - mn * n1;
    m2 = m2 - mn * n2;
    const double V[5] = {r, m0, m1, m2, pr / g1 + 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / r};
#pragma unroll
    for (int k = 0; k < 5; ++k) (c == 0 ? io.R[k][0] : c == 1 ? io.R[k][1] : io.R[k][2])[q] = V[k];
}

template <typename T>
__global__ void k_zero_at5(long n, const int* __restrict__ sel, P5<T> s)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long i = sel[q];
#pragma unroll
    for (int k = 0; k < 5; ++k) s.p[k][i] = T(0);
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

// A stage input: its mesh-point part (FP64, or FP32 in the mixed form) and
// its cut-point parts (FP64).
template <typename TU>
struct StageIn {
    const TU* D[5];
    const double* R[5][3];
};

template <typename TD>
int wall_launch(shx_euler* e, const TD* const D[5], double* const R[5][3])
{
    WallIO<TD> io{};
    io.off[0] = 0;
    for (int c = 0; c < 3; ++c) io.off[c + 1] = io.off[c] + e->nR[c];
    if (io.off[3] == 0) return SHX_OK;
    WallArgs w;
    for (int k = 0; k < 5; ++k) {
        io.D[k] = D[k];
        for (int c = 0; c < 3; ++c) io.R[k][c] = R[k][c];
    }
    for (int c = 0; c < 3; ++c) {
        w.fluid[c] = e->wall_fluid[c];
        w.nrm[c] = e->wall_normal[c];
    }
    k_wall<TD><<<blocks_for(io.off[3]), kT, 0, e->ctx->stream>>>(io, w, e->gamma - 1.0);
    return check(e->ctx, "wall kernel");
}

int wall_field(shx_euler* e, shx_field* U)
{
    const double* D[5];
    double* R[5][3];
    for (int k = 0; k < 5; ++k) {
        D[k] = U[k].D;
        for (int c = 0; c < 3; ++c) R[k][c] = U[k].R[c];
    }
    return wall_launch<double>(e, D, R);
}

// the negated fluxes along dir at the cut points of a stage input
void cut_fluxes(shx_euler* e, const double* const R[5][3], int dir)
{
    Five4 fs{};
    for (int k = 0; k < 5; ++k)
        for (int c = 0; c < 3; ++c) {
            fs.a[k][1 + c] = R[k][c];
            fs.d[k][1 + c] = e->FR[k][c];
        }
    fs.off[1] = 0;
    for (int c = 0; c < 3; ++c) fs.off[2 + c] = fs.off[1 + c] + e->nR[c];
    if (fs.off[4] > 0) k_flux_neg<<<blocks_for(fs.off[4]), kT, 0, e->ctx->stream>>>(fs, dir, e->gamma);
}

template <int DIR, int W, typename FM>
void dflux_launch(shx_euler* e, const shx_derivative* d, bool plus, P5<const typename FM::TU> U,
                  P5<typename FM::TR> R)
{
    hipStream_t st = e->ctx->stream;
    const double g = e->gamma;
    // rows along x: four planes per lane when the grid keeps at least eight
    // workgroups per CU
    const int g4[3] = {(d->n[0] + kDerivPlanes - 1) / kDerivPlanes, d->n[1], d->n[2]};
    const Tiling t4 = make_tiling(g4);
    if (DIR == 0 && t4.nblocks >= 8u * 256u) {
        if (plus) k_euler_dflux_x<W, true, kDerivPlanes, FM><<<t4.nblocks, kBlock, 0, st>>>(t4, d->n[0], d->dev, g, U, R);
        else k_euler_dflux_x<W, false, kDerivPlanes, FM><<<t4.nblocks, kBlock, 0, st>>>(t4, d->n[0], d->dev, g, U, R);
    } else {
        const Tiling t = make_tiling(d->n);
        if (plus) k_euler_dflux<DIR, W, true, FM><<<t.nblocks, kBlock, 0, st>>>(t, d->dev, g, U, R);
        else k_euler_dflux<DIR, W, false, FM><<<t.nblocks, kBlock, 0, st>>>(t, d->dev, g, U, R);
    }
}

template <int W, typename FM>
void list_launch(shx_euler* e, const shx_derivative* d, int dir, bool plus, P5<const typename FM::TU> U,
                 P5<typename FM::TR> R)
{
    if (d->nslow == 0) return;
    P5<const double> FR;
    for (int k = 0; k < 5; ++k) FR.p[k] = e->FR[k][dir];
    const unsigned grid = (unsigned)((d->nslow + 255) / 256);
    hipStream_t st = e->ctx->stream;
    if (plus) k_euler_dflux_list<W, true, FM><<<grid, 256, 0, st>>>(d->nslow, d->slow, d->dev, dir, e->gamma, U, FR, R);
    else k_euler_dflux_list<W, false, FM><<<grid, 256, 0, st>>>(d->nslow, d->slow, d->dev, dir, e->gamma, U, FR, R);
}

// one direction's interior pass (assign or add) for the stencil width and
// direction at run time
template <typename FM>
int dflux_dispatch(shx_euler* e, int dir, bool plus, P5<const typename FM::TU> U, P5<typename FM::TR> R)
{
    const shx_derivative* d = e->d[dir];
    switch (dir * 10 + d->dev.width) {
    case 3: dflux_launch<0, 3, FM>(e, d, plus, U, R); break;
    case 5: dflux_launch<0, 5, FM>(e, d, plus, U, R); break;
    case 13: dflux_launch<1, 3, FM>(e, d, plus, U, R); break;
    case 15: dflux_launch<1, 5, FM>(e, d, plus, U, R); break;
    case 23: dflux_launch<2, 3, FM>(e, d, plus, U, R); break;
    case 25: dflux_launch<2, 5, FM>(e, d, plus, U, R); break;
    default: return fail(e->ctx, SHX_EUNSUPPORTED, "euler: interior width must be 3 or 5");
    }
    return SHX_OK;
}

template <typename FM>
int list_dispatch(shx_euler* e, int dir, bool plus, P5<const typename FM::TU> U, P5<typename FM::TR> R)
{
    const shx_derivative* d = e->d[dir];
    if (d->dev.width == 3) list_launch<3, FM>(e, d, dir, plus, U, R);
    else if (d->dev.width == 5) list_launch<5, FM>(e, d, dir, plus, U, R);
    else return fail(e->ctx, SHX_EUNSUPPORTED, "euler: interior width must be 3 or 5");
    return SHX_OK;
}

template <int DIR, int W, int OP, typename FM, typename TO>
void stage_pass_launch(shx_euler* e, P5<const typename FM::TU> U, P5<typename FM::TR> R, const StageOut<TO>& so)
{
    const shx_derivative* d = e->d[DIR];
    const Tiling t = make_tiling(d->n);
    k_euler_dflux_stage<DIR, W, OP, FM, TO>
        <<<t.nblocks, kBlock, 0, e->ctx->stream>>>(t, d->dev, e->gamma, U, R, so, e->face_mask);
}

template <int OP, typename FM, typename TO>
int stage_pass_dispatch(shx_euler* e, int dir, P5<const typename FM::TU> U, P5<typename FM::TR> R,
                        const StageOut<TO>& so)
{
    switch (dir * 10 + e->d[dir]->dev.width) {
    case 13: stage_pass_launch<1, 3, OP, FM, TO>(e, U, R, so); break;
    case 15: stage_pass_launch<1, 5, OP, FM, TO>(e, U, R, so); break;
    case 23: stage_pass_launch<2, 3, OP, FM, TO>(e, U, R, so); break;
    case 25: stage_pass_launch<2, 5, OP, FM, TO>(e, U, R, so); break;
    default: return fail(e->ctx, SHX_EUNSUPPORTED, "euler: interior width must be 3 or 5");
    }
    return SHX_OK;
}

// directions with an operator, in order
int present_dirs(const shx_euler* e, int (&dirs)[3])
{
    int n = 0;
    for (int c = 0; c < 3; ++c)
        if (e->n[c] >= 2 && e->d[c]) dirs[n++] = c;
    return n;
}

// rhs of the stage input `in` into the partial-rhs buffer R for every
// direction but the last, with the last direction's listed rows added
// (the last direction's interior pass is the caller's)
template <typename FM>
int partial_rhs(shx_euler* e, const StageIn<typename FM::TU>& in, P5<typename FM::TR> R, int (&dirs)[3], int nd)
{
    hipStream_t st = e->ctx->stream;
    if (e->nnorow) k_zero_at5<typename FM::TR><<<blocks_for(e->nnorow), kT, 0, st>>>(e->nnorow, e->norow, R);
    int rc = check(e->ctx, "rhs fill");
    P5<const typename FM::TU> U;
    for (int k = 0; k < 5; ++k) U.p[k] = in.D[k];
    for (int a = 0; a < nd && rc == SHX_OK; ++a) {
        const int dir = dirs[a];
        cut_fluxes(e, in.R, dir);
        if (a + 1 < nd) rc = dflux_dispatch<FM>(e, dir, a > 0, U, R);
        if (rc == SHX_OK) rc = list_dispatch<FM>(e, dir, a > 0, U, R);
        if (rc == SHX_OK) rc = check(e->ctx, "euler flux-derivative kernels");
    }
    return rc;
}

// the whole rhs into rhs (FP64 form), as shx_euler_rhs defines it
int rhs_launch(shx_euler* e, const shx_field* U, shx_field* rhs)
{
    hipStream_t st = e->ctx->stream;
    // the cut points of the rhs are zero
    {
        Five4 z = five4(e, rhs, nullptr, nullptr);
        for (int k = 0; k < 5; ++k) z.d[k][0] = nullptr;
        z.off[1] = 0;
        for (int c = 0; c < 3; ++c) z.off[2 + c] = z.off[1 + c] + e->nR[c];
        if (z.off[4] > 0) k_slot5<kSlotZero><<<dim3(blocks_for(z.off[4]), 5), kT, 0, st>>>(z, 0.0);
    }
    StageIn<double> in;
    P5<double> R;
    for (int k = 0; k < 5; ++k) {
        in.D[k] = U[k].D;
        for (int c = 0; c < 3; ++c) in.R[k][c] = U[k].R[c];
        R.p[k] = rhs[k].D;
    }
    int dirs[3];
    const int nd = present_dirs(e, dirs);
    int rc = partial_rhs<F64>(e, in, R, dirs, nd);
    if (rc == SHX_OK && nd > 0) {
        P5<const double> Uc;
        for (int k = 0; k < 5; ++k) Uc.p[k] = U[k].D;
        rc = dflux_dispatch<F64>(e, dirs[nd - 1], nd > 1, Uc, R);
    }
    if (rc != SHX_OK) return rc;
    if (e->nfaces) k_zero_at5<double><<<blocks_for(e->nfaces), kT, 0, st>>>(e->nfaces, e->faces, R);
    return check(e->ctx, "rhs finish kernels");
}

// One stage: partial rhs of `in` into R, then the last direction's pass
// with the slot operations into so.out, then the wall on the new input.
template <int OP, typename FM, typename TO>
int stage_launch(shx_euler* e, const StageIn<typename FM::TU>& in, P5<typename FM::TR> R, const StageOut<TO>& so,
                 double* const outR[5][3])
{
    int dirs[3];
    const int nd = present_dirs(e, dirs);
    int rc = partial_rhs<FM>(e, in, R, dirs, nd);
    P5<const typename FM::TU> U;
    for (int k = 0; k < 5; ++k) U.p[k] = in.D[k];
    if (rc == SHX_OK) rc = stage_pass_dispatch<OP, FM, TO>(e, dirs[nd - 1], U, R, so);
    if (rc == SHX_OK) rc = check(e->ctx, "euler stage kernel");
    const TO* D[5];
    for (int k = 0; k < 5; ++k) D[k] = so.out[k];
    if (rc == SHX_OK) rc = wall_launch<TO>(e, D, outR);
    return rc;
}

static const double kRki[4] = {0.0, 0.5, 0.5, 1.0};
static const double kRkf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

// One RK4 step, enqueued.  The stage inputs alternate between two buffers
// (FP64: srhs and out; mixed: the two FP32 buffers), each stage's partial
// rhs going to the buffer that is not its input; the cut points of every
// stage input live in out.R; the result is written to out.
int step_launch(shx_euler* e, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs, double dt)
{
    double* outR[5][3];
    StageIn<double> in0;
    for (int k = 0; k < 5; ++k) {
        in0.D[k] = u0[k].D;
        for (int c = 0; c < 3; ++c) {
            outR[k][c] = out[k].R[c];
            in0.R[k][c] = u0[k].R[c];
        }
    }
    auto stage_out = [&](auto* dst, int i) {
        using TO = std::remove_pointer_t<std::remove_pointer_t<decltype(dst)>>;
        StageOut<TO> so;
        for (int k = 0; k < 5; ++k) {
            so.u0[k] = u0[k].D;
            so.rk[k] = rk[k].D;
            so.out[k] = dst[k];
        }
        so.c1 = dt * kRkf[i];
        so.c2 = i < 3 ? dt * kRki[i + 1] : 1.0;
        return so;
    };
    int rc = SHX_OK;
    if (e->precision == SHX_F64) {
        double* buf[2][5];
        for (int k = 0; k < 5; ++k) {
            buf[0][k] = srhs[k].D;
            buf[1][k] = out[k].D;
        }
        // stage i: partial rhs and next input in buf[i % 2], input in buf[(i + 1) % 2] (u0 at stage 0)
        for (int i = 0; i < 4 && rc == SHX_OK; ++i) {
            const int w = i % 2;
            StageIn<double> in = in0;
            if (i > 0)
                for (int k = 0; k < 5; ++k) {
                    in.D[k] = buf[1 - w][k];
                    for (int c = 0; c < 3; ++c) in.R[k][c] = outR[k][c];
                }
            P5<double> R;
            for (int k = 0; k < 5; ++k) R.p[k] = buf[w][k];
            // the result goes to out: the last stage (i = 3) has w = 1
            const auto so = stage_out(buf[w], i);
            if (i == 0) rc = stage_launch<kStepFirst, F64, double>(e, in, R, so, outR);
            else if (i < 3) rc = stage_launch<kStepMid, F64, double>(e, in, R, so, outR);
            else rc = stage_launch<kStepLast, F64, double>(e, in, R, so, outR);
        }
        return rc;
    }
    // mixed form
    auto& fb = e->mixF;
    for (int i = 0; i < 4 && rc == SHX_OK; ++i) {
        const int w = i % 2;
        P5<float> R;
        for (int k = 0; k < 5; ++k) R.p[k] = fb[w][k];
        if (i == 0) {
            rc = stage_launch<kStepFirst, Mixed0, float>(e, in0, R, stage_out(fb[w], i), outR);
            continue;
        }
        StageIn<float> in;
        for (int k = 0; k < 5; ++k) {
            in.D[k] = fb[1 - w][k];
            for (int c = 0; c < 3; ++c) in.R[k][c] = outR[k][c];
        }
        if (i < 3) {
            rc = stage_launch<kStepMid, Mixed, float>(e, in, R, stage_out(fb[w], i), outR);
        } else {
            double* res[5];
            for (int k = 0; k < 5; ++k) res[k] = out[k].D;
            rc = stage_launch<kStepLast, Mixed, double>(e, in, R, stage_out(res, i), outR);
        }
    }
    return rc;
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

void free_mixed(shx_euler* e)
{
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 5; ++k) {
            hipFree(e->mixF[b][k]);
            e->mixF[b][k] = nullptr;
        }
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
    {
        int np = 0;
        for (int c = 0; c < 3; ++c) np += (d[c] && any->n[c] >= 2) ? 1 : 0;
        if (np < 2)
            return fail(ctx, SHX_EUNSUPPORTED,
                        "euler: needs at least two directions (the last one's pass completes each RK stage)");
    }
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
                e->face_mask |= 1u << (2 * a + side);
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
    free_mixed(e);
    delete e;
    return SHX_OK;
}

int shx_euler_set_precision(shx_euler* e, int precision)
{
    if (!e || (precision != SHX_F64 && precision != SHX_F32_COMP)) return SHX_EINVAL;
    if (precision == SHX_F32_COMP && !e->mixF[0][0]) {
        hipError_t err = hipSuccess;
        for (int b = 0; b < 2 && err == hipSuccess; ++b)
            for (int k = 0; k < 5 && err == hipSuccess; ++k)
                err = hipMalloc((void**)&e->mixF[b][k], sizeof(float) * (size_t)e->N);
        if (err != hipSuccess) {
            free_mixed(e);
            return hip_fail(e->ctx, err, "euler mixed-form buffers");
        }
    }
    e->precision = precision;
    return SHX_OK;
}

int shx_euler_wall(shx_euler* e, shx_field* U)
{
    if (!e || !U) return SHX_EINVAL;
    const int rc = wall_field(e, U);
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
    const int rc = step_launch(e, u0, out, rk, srhs, dt);
    return rc == SHX_OK ? sync(e->ctx) : rc;
}

int shx_time_euler_steps(shx_euler* e, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs, double dt,
                         int nsteps, double* ms_per_step)
{
    if (!e || !u0 || !out || !rk || !srhs || nsteps < 1 || !ms_per_step) return SHX_EINVAL;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return hip_fail(e->ctx, hipGetLastError(), "event");
    if (hipEventCreate(&b) != hipSuccess) {
        hipEventDestroy(a);
        return hip_fail(e->ctx, hipGetLastError(), "event");
    }
    int rc = sync(e->ctx);
    hipEventRecord(a, e->ctx->stream);
    // u0 <- out after each step (simulation_cycle.cpp:61): the two swap
    // roles from step to step, and one copy at the end leaves the result
    // in both
    shx_field* x = u0;
    shx_field* y = out;
    for (int s = 0; s < nsteps && rc == SHX_OK; ++s) {
        rc = step_launch(e, x, y, rk, srhs, dt);
        shx_field* t = x;
        x = y;
        y = t;
    }
    if (rc == SHX_OK) {
        // x holds the result
        if (x == u0) slot5<kSlotCopy>(e, out, u0, 0.0, nullptr);
        else slot5<kSlotCopy>(e, u0, out, 0.0, nullptr);
        rc = check(e->ctx, "euler copy");
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
