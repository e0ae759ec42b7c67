// shx_adjoint.hip -- transposed derivative applies and the discrete adjoint of
// the RK4 step (no reference counterpart: the reference has no adjoint).
//
// Transpose of one direction's derivative, du.D = O u.D + B u.R[dir]:
//   * a column q is *regular* when the rows at positions q - hw .. q + hw of
//     its line are all interior (circulant) rows and no closure row reads q;
//     then (O^T y)[q] = sum_m coef[m] y[q - (m - hw) stride], the reversed
//     stencil, computed in place with no index data;
//   * every other column (walls, bodies, solid points, columns read by
//     closures) is listed with its entries, rows ascending (CSR built once on
//     the host from the line segments);
//   * B^T: per cut point of the direction, its entries, rows ascending.
// The RK4 adjoint recomputes the step's stage inputs from the step's input,
// then runs the stages backwards (see shx_rk4_adjoint in shoccs_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "shx_internal.hpp"
#include "shx_point.hpp"
#include "shx_state.hpp"
#include "shx_sys.hpp"
#include "shx_system_state.hpp"

using namespace shx;

struct shx_deriv_T {
    long N = 0, nR = 0;
    unsigned* irr_mask = nullptr;  // bit per point: column needs its listed entries
    unsigned* int_mask = nullptr;  // bit per point: the point's row is a circulant row
    long nirr = 0;
    int* irr_idx = nullptr;  // irregular columns, ascending
    int* irr_ptr = nullptr;
    int* irr_row = nullptr;
    double* irr_w = nullptr;
    int* b_ptr = nullptr;  // per cut point of the direction
    int* b_row = nullptr;
    double* b_w = nullptr;
};

// Adjoint workspace of a system: stage inputs w1..w3, the stage-RHS scratch,
// the adjoint registers and the per-direction weighted adjoints.
struct shx_adjoint_ws {
    shx_field w[3];
    shx_field k, g, kb, wb;
    double* ybar[3] = {nullptr, nullptr, nullptr};
    double* part = nullptr;  // coefficient-gradient partials
    double* gacc = nullptr;  // coefficient gradient accumulated on the device, 3 x 9
    long npart = 0;
};

namespace {

constexpr int kT = 256;
inline unsigned blocks_for(long n) { return (unsigned)((n + kT - 1) / kT); }

int check(shx_ctx* ctx, const char* what)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHX_OK : hip_fail(ctx, e, what);
}

__device__ __forceinline__ bool bit(const unsigned* m, long q) { return (m[q >> 5] >> (q & 31)) & 1u; }

template <int W>
__global__ void k_dT_regular(long N, long st, const double* __restrict__ coef,
                             const unsigned* __restrict__ irr, const double* __restrict__ y,
                             double* __restrict__ x, int plus)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N || bit(irr, q)) return;
    constexpr int hw = W / 2;
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < W; ++m) s += coef[m] * y[q - (long)(m - hw) * st];
    x[q] = plus ? x[q] + s : s;
}

// Listed entries: x[idx[e]] (= | +=) sum over ptr[e] .. ptr[e+1] of w y[row]
// (idx == nullptr: x[e]).
__global__ void k_dT_list(long n, const int* __restrict__ idx, const int* __restrict__ ptr,
                          const int* __restrict__ row, const double* __restrict__ w,
                          const double* __restrict__ y, double* __restrict__ x, int plus)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = 0.0;
    for (int k = ptr[e]; k < ptr[e + 1]; ++k) s += w[k] * y[row[k]];
    const long c = idx ? (long)idx[e] : e;
    x[c] = plus ? x[c] + s : s;
}

__global__ void k_fill0(long n, double* __restrict__ x)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] = 0.0;
}

// out = a + c b
__global__ void k_lc(long n, const double* __restrict__ a, const double* __restrict__ b, double c,
                     double* __restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = a[e] + c * b[e];
}

// out = c a
__global__ void k_scale(long n, double c, const double* __restrict__ a, double* __restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = c * a[e];
}

// out = ca a + cb b
__global__ void k_lc2(long n, double ca, const double* __restrict__ a, double cb,
                      const double* __restrict__ b, double* __restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = ca * a[e] + cb * b[e];
}

// x += y
__global__ void k_add(long n, const double* __restrict__ y, double* __restrict__ x)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] += y[e];
}

__global__ void k_zero_list(long n, const int* __restrict__ idx, double* __restrict__ x)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[idx[e]] = 0.0;
}

__global__ void k_zero_flag(long n, const signed char* __restrict__ flag, double* __restrict__ x)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n && flag[e]) x[e] = 0.0;
}

// The rhs's derivative weights applied to an adjoint: for direction c,
// ybar[c] = d rhs / d (D_c u) * kbar, zero on Dirichlet faces.  Scalar wave:
// gG_c (scalar_wave.cpp:28-33 and the constant-velocity form); heat: k, the
// same for every direction (heat.cpp:119-162).
template <bool HEAT>
__global__ void k_adj_weights(DevSys s, long N, const double* __restrict__ kbar, double* __restrict__ y0,
                              double* __restrict__ y1, double* __restrict__ y2)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    const bool face = on_dirichlet_face(s, i, j, k);
    const double kb = kbar[idx];
    if (HEAT) {
        y0[idx] = face ? 0.0 : s.k * kb;
        return;
    }
    double g[3] = {0.0, 0.0, 0.0};
    if (!face) {
        const double x = s.x[i], y = s.y[j], z = s.z[k];
        for (int c = 0; c < 3; ++c) g[c] = wave_neg_g(s, c, x, y, z);
    }
    if (y0) y0[idx] = g[0] * kb;
    if (y1) y1[idx] = g[1] * kb;
    if (y2) y2[idx] = g[2] * kb;
}

// Per-workgroup partial sums, over the points whose row is a circulant row,
// of ybar[p] w[p + (m - hw) stride], m < W.
template <int W>
__global__ __launch_bounds__(kT) void k_coef_part(long N, long st, const unsigned* __restrict__ introw,
                                                  const double* __restrict__ ybar,
                                                  const double* __restrict__ w, double* __restrict__ part)
{
    constexpr int hw = W / 2;
    __shared__ double red[W][kT];
    double acc[W];
#pragma unroll
    for (int m = 0; m < W; ++m) acc[m] = 0.0;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
        if (!bit(introw, p)) continue;
        const double yb = ybar[p];
#pragma unroll
        for (int m = 0; m < W; ++m) acc[m] += yb * w[p + (long)(m - hw) * st];
    }
#pragma unroll
    for (int m = 0; m < W; ++m) red[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int h = kT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
#pragma unroll
            for (int m = 0; m < W; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x < W) part[(long)blockIdx.x * W + threadIdx.x] = red[threadIdx.x][0];
}

// acc[m] += (sum over workgroups of part[b][m], from 0.0 in workgroup order).
// The partials are staged through LDS a chunk at a time by the whole
// workgroup (coalesced loads in flight together); lane m then adds its
// column of the chunk in order, so the sum is the sequential one.
constexpr int kReduceChunk = 256;
__global__ __launch_bounds__(kT) void k_coef_reduce(int nb, int W, const double* __restrict__ part,
                                                    double* __restrict__ acc)
{
    __shared__ double buf[kReduceChunk * 9];
    const int m = threadIdx.x;
    double s = 0.0;
    for (int b0 = 0; b0 < nb; b0 += kReduceChunk) {
        const int nc = min(kReduceChunk, nb - b0);
        for (int e = threadIdx.x; e < nc * W; e += blockDim.x) buf[e] = part[(long)b0 * W + e];
        __syncthreads();
        if (m < W)
            for (int b = 0; b < nc; ++b) s += buf[b * W + m];
        __syncthreads();
    }
    if (m < W) acc[m] += s;
}

// ------------------------------------------------------------ transpose build
long line_base(const shx_derivative* d, int pair)
{
    const int nfast = d->dev.nfast;
    const int s = pair / nfast, f = pair % nfast;
    if (d->dir == 0) return (long)s * d->n[2] + f;
    if (d->dir == 1) return (long)s * d->n[1] * d->n[2] + f;
    return ((long)s * d->n[1] + f) * d->n[2];
}

template <typename T>
int up(shx_ctx* ctx, T** dst, const std::vector<T>& v)
{
    *dst = nullptr;
    if (v.empty()) return SHX_OK;
    return upload(ctx, dst, v.data(), v.size());
}

int build_transpose(shx_derivative* d)
{
    if (d->T) return SHX_OK;
    shx_ctx* ctx = d->ctx;
    const long N = (long)d->n[0] * d->n[1] * d->n[2];
    const long st = d->dev.stride;
    const int nq = d->n[d->dir];
    const int W = d->dev.width, hw = W / 2;
    const int nslow = d->dir == 0 ? d->n[1] : d->n[0];
    const long npairs = (long)nslow * d->dev.nfast;
    const long nR = d->nR[d->dir];
    if (N >= 0x7fffffffL) return fail(ctx, SHX_EUNSUPPORTED, "transpose: mesh too large");
    // row type per point: 0 none, 1 circulant, 2 closure
    std::vector<unsigned char> rt((size_t)N, 0);
    for (const shx_line_block& b : d->h_blocks) {
        const long base = line_base(d, b.pair);
        const int total = b.nl + b.ni + b.nr;
        for (int lr = 0; lr < total; ++lr)
            rt[(size_t)(base + (long)(b.lo + lr) * st)] = (lr >= b.nl && lr < b.nl + b.ni) ? 1 : 2;
    }
    std::vector<unsigned char> irr((size_t)N, 0);
    for (long pair = 0; pair < npairs; ++pair) {
        const long base = line_base(d, (int)pair);
        for (int qp = 0; qp < nq; ++qp) {
            bool reg = true;
            for (int m = 0; m < W && reg; ++m) {
                const int pp = qp + hw - m;
                reg = pp >= 0 && pp < nq && rt[(size_t)(base + (long)pp * st)] == 1;
            }
            irr[(size_t)(base + (long)qp * st)] = !reg;
        }
    }
    auto closure = [&](const shx_line_block& b, int lr, int& col0, int& ncols, const double*& a) {
        if (lr < b.nl) {
            col0 = b.lcol0;
            ncols = b.lcols;
            a = d->h_pool.data() + b.lcoef + (long)lr * b.lcols;
        } else {
            const int rr = lr - b.nl - b.ni;
            col0 = b.rcol0;
            ncols = b.rcols;
            a = d->h_pool.data() + b.rcoef + (long)rr * b.rcols;
        }
    };
    for (const shx_line_block& b : d->h_blocks) {
        const long base = line_base(d, b.pair);
        const int total = b.nl + b.ni + b.nr;
        for (int lr = 0; lr < total; ++lr) {
            if (lr >= b.nl && lr < b.nl + b.ni) continue;
            int col0, ncols;
            const double* a;
            closure(b, lr, col0, ncols, a);
            for (int jj = 0; jj < ncols; ++jj) irr[(size_t)(base + (long)(col0 + jj) * st)] = 1;
        }
    }
    struct Ent {
        long col;
        long row;
        double w;
    };
    std::vector<Ent> ent, bent;
    for (const shx_line_block& b : d->h_blocks) {
        const long base = line_base(d, b.pair);
        const int total = b.nl + b.ni + b.nr;
        for (int lr = 0; lr < total; ++lr) {
            const long p = base + (long)(b.lo + lr) * st;
            if (lr >= b.nl && lr < b.nl + b.ni) {
                for (int jj = 0; jj < W; ++jj) {
                    const long c = p + (long)(jj - hw) * st;
                    if (irr[(size_t)c]) ent.push_back({c, p, d->h_coef[jj]});
                }
            } else {
                int col0, ncols;
                const double* a;
                closure(b, lr, col0, ncols, a);
                for (int jj = 0; jj < ncols; ++jj) ent.push_back({base + (long)(col0 + jj) * st, p, a[jj]});
            }
            if (b.lobj >= 0 && lr < b.nl) bent.push_back({b.lobj, p, d->h_pool[(size_t)b.lbw + lr]});
            if (b.robj >= 0 && lr >= total - b.nr)
                bent.push_back({b.robj, p, d->h_pool[(size_t)b.rbw + (lr - (total - b.nr))]});
        }
    }
    auto by_col_row = [](const Ent& x, const Ent& y) { return x.col != y.col ? x.col < y.col : x.row < y.row; };
    std::stable_sort(ent.begin(), ent.end(), by_col_row);
    std::stable_sort(bent.begin(), bent.end(), by_col_row);
    if ((long)ent.size() >= 0x7fffffffL || (long)bent.size() >= 0x7fffffffL)
        return fail(ctx, SHX_EUNSUPPORTED, "transpose: too many entries");

    std::vector<int> irr_idx, irr_ptr{0}, irr_row;
    std::vector<double> irr_w;
    size_t e = 0;
    for (long q = 0; q < N; ++q) {
        if (!irr[(size_t)q]) continue;
        irr_idx.push_back((int)q);
        for (; e < ent.size() && ent[e].col == q; ++e) {
            irr_row.push_back((int)ent[e].row);
            irr_w.push_back(ent[e].w);
        }
        irr_ptr.push_back((int)irr_row.size());
    }
    std::vector<int> b_ptr{0}, b_row;
    std::vector<double> b_w;
    e = 0;
    for (long o = 0; o < nR; ++o) {
        for (; e < bent.size() && bent[e].col == o; ++e) {
            b_row.push_back((int)bent[e].row);
            b_w.push_back(bent[e].w);
        }
        b_ptr.push_back((int)b_row.size());
    }
    std::vector<unsigned> irr_mask((size_t)(N + 31) / 32, 0u), int_mask((size_t)(N + 31) / 32, 0u);
    for (long q = 0; q < N; ++q) {
        if (irr[(size_t)q]) irr_mask[(size_t)(q >> 5)] |= 1u << (q & 31);
        if (rt[(size_t)q] == 1) int_mask[(size_t)(q >> 5)] |= 1u << (q & 31);
    }

    auto* T = new (std::nothrow) shx_deriv_T();
    if (!T) return SHX_ENOMEM;
    T->N = N;
    T->nR = nR;
    T->nirr = (long)irr_idx.size();
    int rc = up(ctx, &T->irr_mask, irr_mask);
    if (rc == SHX_OK) rc = up(ctx, &T->int_mask, int_mask);
    if (rc == SHX_OK) rc = up(ctx, &T->irr_idx, irr_idx);
    if (rc == SHX_OK) rc = up(ctx, &T->irr_ptr, irr_ptr);
    if (rc == SHX_OK) rc = up(ctx, &T->irr_row, irr_row);
    if (rc == SHX_OK) rc = up(ctx, &T->irr_w, irr_w);
    if (rc == SHX_OK) rc = up(ctx, &T->b_ptr, b_ptr);
    if (rc == SHX_OK) rc = up(ctx, &T->b_row, b_row);
    if (rc == SHX_OK) rc = up(ctx, &T->b_w, b_w);
    if (rc != SHX_OK) {
        destroy_transpose(T);
        return rc;
    }
    d->T = T;
    return SHX_OK;
}

// x = O^T y (D part) and xR = B^T y (R part of the direction), or += when plus
int apply_T(shx_derivative* d, const double* y, double* x, double* xR, bool plus)
{
    int rc = build_transpose(d);
    if (rc != SHX_OK) return rc;
    shx_deriv_T* T = d->T;
    hipStream_t st = d->ctx->stream;
    const int p = plus ? 1 : 0;
    switch (d->dev.width) {
    case 3: k_dT_regular<3><<<blocks_for(T->N), kT, 0, st>>>(T->N, d->dev.stride, d->coef, T->irr_mask, y, x, p); break;
    case 5: k_dT_regular<5><<<blocks_for(T->N), kT, 0, st>>>(T->N, d->dev.stride, d->coef, T->irr_mask, y, x, p); break;
    case 7: k_dT_regular<7><<<blocks_for(T->N), kT, 0, st>>>(T->N, d->dev.stride, d->coef, T->irr_mask, y, x, p); break;
    case 9: k_dT_regular<9><<<blocks_for(T->N), kT, 0, st>>>(T->N, d->dev.stride, d->coef, T->irr_mask, y, x, p); break;
    default: return fail(d->ctx, SHX_EUNSUPPORTED, "transpose: interior width must be 3, 5, 7 or 9");
    }
    if (T->nirr > 0)
        k_dT_list<<<blocks_for(T->nirr), kT, 0, st>>>(T->nirr, T->irr_idx, T->irr_ptr, T->irr_row, T->irr_w, y, x, p);
    if (T->nR > 0 && xR)
        k_dT_list<<<blocks_for(T->nR), kT, 0, st>>>(T->nR, nullptr, T->b_ptr, T->b_row, T->b_w, y, xR, p);
    return check(d->ctx, "transposed derivative");
}

// ----------------------------------------------------------- system adjoint
void free_field(shx_field& f)
{
    hipFree(f.D);
    for (int c = 0; c < 3; ++c) hipFree(f.R[c]);
    std::memset(&f, 0, sizeof f);
}

int alloc_field(shx_system* s, shx_field& f)
{
    std::memset(&f, 0, sizeof f);
    hipError_t e = hipMalloc((void**)&f.D, sizeof(double) * (size_t)std::max<long>(s->N, 1));
    for (int c = 0; c < 3 && e == hipSuccess; ++c)
        e = hipMalloc((void**)&f.R[c], sizeof(double) * (size_t)std::max<long>(s->nR[c], 1));
    return e == hipSuccess ? SHX_OK : hip_fail(s->ctx, e, "adjoint workspace");
}

int ensure_ws(shx_system* s)
{
    if (s->adj) return SHX_OK;
    auto* w = new (std::nothrow) shx_adjoint_ws();
    if (!w) return SHX_ENOMEM;
    s->adj = w;
    int rc = SHX_OK;
    for (int i = 0; i < 3 && rc == SHX_OK; ++i) rc = alloc_field(s, w->w[i]);
    for (shx_field* f : {&w->k, &w->g, &w->kb, &w->wb})
        if (rc == SHX_OK) rc = alloc_field(s, *f);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
        hipError_t e = hipMalloc((void**)&w->ybar[c], sizeof(double) * (size_t)std::max<long>(s->N, 1));
        if (e != hipSuccess) rc = hip_fail(s->ctx, e, "adjoint workspace");
    }
    if (rc == SHX_OK) {
        w->npart = std::min<long>(2048, std::max<long>(1, (s->N + kT - 1) / kT));
        hipError_t e = hipMalloc((void**)&w->part, sizeof(double) * (size_t)w->npart * 9);
        if (e == hipSuccess) e = hipMalloc((void**)&w->gacc, sizeof(double) * 27);
        if (e != hipSuccess) rc = hip_fail(s->ctx, e, "adjoint workspace");
    }
    if (rc != SHX_OK) {
        destroy_adjoint_ws(w);
        s->adj = nullptr;
    }
    return rc;
}

// x = a + c b over every part of the fields
void field_lc(shx_system* s, const shx_field& a, const shx_field& b, double c, shx_field& out)
{
    hipStream_t st = s->ctx->stream;
    k_lc<<<blocks_for(s->N), kT, 0, st>>>(s->N, a.D, b.D, c, out.D);
    for (int r = 0; r < 3; ++r)
        if (s->nR[r]) k_lc<<<blocks_for(s->nR[r]), kT, 0, st>>>(s->nR[r], a.R[r], b.R[r], c, out.R[r]);
}

void field_lc2(shx_system* s, double ca, const shx_field& a, double cb, const shx_field& b, shx_field& out)
{
    hipStream_t st = s->ctx->stream;
    k_lc2<<<blocks_for(s->N), kT, 0, st>>>(s->N, ca, a.D, cb, b.D, out.D);
    for (int r = 0; r < 3; ++r)
        if (s->nR[r]) k_lc2<<<blocks_for(s->nR[r]), kT, 0, st>>>(s->nR[r], ca, a.R[r], cb, b.R[r], out.R[r]);
}

void field_scale(shx_system* s, double c, const shx_field& a, shx_field& out)
{
    hipStream_t st = s->ctx->stream;
    k_scale<<<blocks_for(s->N), kT, 0, st>>>(s->N, c, a.D, out.D);
    for (int r = 0; r < 3; ++r)
        if (s->nR[r]) k_scale<<<blocks_for(s->nR[r]), kT, 0, st>>>(s->nR[r], c, a.R[r], out.R[r]);
}

void field_add(shx_system* s, const shx_field& y, shx_field& x)
{
    hipStream_t st = s->ctx->stream;
    k_add<<<blocks_for(s->N), kT, 0, st>>>(s->N, y.D, x.D);
    for (int r = 0; r < 3; ++r)
        if (s->nR[r]) k_add<<<blocks_for(s->nR[r]), kT, 0, st>>>(s->nR[r], y.R[r], x.R[r]);
}

// P: zero the entries update_boundary overwrites (Dirichlet grid faces of D,
// Dirichlet cut points)
void project(shx_system* s, shx_field& f)
{
    hipStream_t st = s->ctx->stream;
    if (s->nfaces) k_zero_list<<<blocks_for(s->nfaces), kT, 0, st>>>(s->nfaces, s->faces, f.D);
    for (int r = 0; r < 3; ++r)
        if (s->nR[r]) k_zero_flag<<<blocks_for(s->nR[r]), kT, 0, st>>>(s->nR[r], s->rdir[r], f.R[r]);
}

// wbar = (d rhs / d u)^T kbar, and, when gacc (device, 3 x 9) is given, the
// stage's coefficient-gradient terms with stage input w added to
// gacc[9 c + m]
int rhs_T(shx_system* s, const shx_field& kbar, shx_field& wbar, const shx_field* w, double* gacc)
{
    shx_adjoint_ws* ws = s->adj;
    hipStream_t st = s->ctx->stream;
    const bool heat = s->kind == SHX_SYS_HEAT;
    if (heat)
        k_adj_weights<true><<<blocks_for(s->N), kT, 0, st>>>(s->dev, s->N, kbar.D, ws->ybar[0], nullptr, nullptr);
    else
        k_adj_weights<false><<<blocks_for(s->N), kT, 0, st>>>(s->dev, s->N, kbar.D, s->d[0] ? ws->ybar[0] : nullptr,
                                                            s->d[1] ? ws->ybar[1] : nullptr,
                                                            s->d[2] ? ws->ybar[2] : nullptr);
    int rc = check(s->ctx, "adjoint weights");
    bool first = true;
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
        if (s->nR[c]) k_fill0<<<blocks_for(s->nR[c]), kT, 0, st>>>(s->nR[c], wbar.R[c]);
        shx_derivative* d = s->d[c];
        if (!d || !d->dev.present) continue;
        const double* y = heat ? ws->ybar[0] : ws->ybar[c];
        rc = apply_T(d, y, wbar.D, wbar.R[c], !first);
        first = false;
        if (rc != SHX_OK || !gacc || !w) continue;
        const int W = d->dev.width;
        const int nb = (int)ws->npart;
        switch (W) {
        case 3: k_coef_part<3><<<nb, kT, 0, st>>>(s->N, d->dev.stride, d->T->int_mask, y, w->D, ws->part); break;
        case 5: k_coef_part<5><<<nb, kT, 0, st>>>(s->N, d->dev.stride, d->T->int_mask, y, w->D, ws->part); break;
        default: return fail(s->ctx, SHX_EUNSUPPORTED, "adjoint: coefficient gradient for widths 3 and 5");
        }
        k_coef_reduce<<<1, kT, 0, st>>>(nb, W, ws->part, gacc + 9 * c);
        rc = check(s->ctx, "adjoint: coefficient gradient");
    }
    if (rc == SHX_OK && first) {  // no direction: the rhs does not depend on u
        k_fill0<<<blocks_for(s->N), kT, 0, st>>>(s->N, wbar.D);
        rc = check(s->ctx, "adjoint");
    }
    return rc;
}

constexpr double kRki[4] = {0.0, 0.5, 0.5, 1.0};
constexpr double kRkf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

}  // namespace

namespace shx {

void destroy_transpose(shx_deriv_T* T)
{
    if (!T) return;
    hipFree(T->irr_mask);
    hipFree(T->int_mask);
    hipFree(T->irr_idx);
    hipFree(T->irr_ptr);
    hipFree(T->irr_row);
    hipFree(T->irr_w);
    hipFree(T->b_ptr);
    hipFree(T->b_row);
    hipFree(T->b_w);
    delete T;
}

void destroy_adjoint_ws(shx_adjoint_ws* w)
{
    if (!w) return;
    for (auto& f : w->w) free_field(f);
    free_field(w->k);
    free_field(w->g);
    free_field(w->kb);
    free_field(w->wb);
    for (auto& y : w->ybar) hipFree(y);
    hipFree(w->part);
    hipFree(w->gacc);
    delete w;
}

}  // namespace shx

// ------------------------------------------------------------------------ C ABI
extern "C" {

int shx_derivative_apply_T(shx_derivative* d, const shx_field* dbar, shx_field* ubar, int accumulate)
{
    if (!d || !dbar || !ubar || !dbar->D || !ubar->D) return SHX_EINVAL;
    if (d->nR[d->dir] > 0 && !ubar->R[d->dir]) return SHX_EINVAL;
    if (d->nrrow > 0)
        return fail(d->ctx, SHX_EUNSUPPORTED,
                    "derivative transpose: operators with cut-point rows (non-Dirichlet bodies) are not supported");
    if (!d->dev.present) {
        if (!accumulate) {
            const long N = (long)d->n[0] * d->n[1] * d->n[2];
            k_fill0<<<blocks_for(N), kT, 0, d->ctx->stream>>>(N, ubar->D);
        }
        const int rc = check(d->ctx, "transposed derivative");
        return rc == SHX_OK ? sync(d->ctx) : rc;
    }
    const int rc = apply_T(d, dbar->D, ubar->D, ubar->R[d->dir], accumulate != 0);
    return rc == SHX_OK ? sync(d->ctx) : rc;
}

int shx_system_rhs_T(shx_system* s, const shx_field* kbar, shx_field* wbar)
{
    if (!s || !kbar || !wbar) return SHX_EINVAL;
    if (s->general || s->hm)
        return fail(s->ctx, SHX_EUNSUPPORTED,
                    "adjoint: floating/Neumann bodies, Neumann grid faces and host-evaluated solutions are not supported");
    int rc = ensure_ws(s);
    if (rc == SHX_OK) rc = rhs_T(s, *kbar, *wbar, nullptr, nullptr);
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_rk4_adjoint(shx_system* s, const shx_field* u, shx_field* ubar, double t, double dt, double* coef_grad)
{
    if (!s || !u || !ubar) return SHX_EINVAL;
    if (s->general || s->hm)
        return fail(s->ctx, SHX_EUNSUPPORTED,
                    "adjoint: floating/Neumann bodies, Neumann grid faces and host-evaluated solutions are not supported");
    int rc = ensure_ws(s);
    if (rc != SHX_OK) return rc;
    shx_adjoint_ws* ws = s->adj;
    // forward stage inputs (rk4.cpp:25-41): w0 = u, w_s = BU(u + dt rki[s] k_{s-1}, t + dt rki[s])
    const shx_field* w[4] = {u, &ws->w[0], &ws->w[1], &ws->w[2]};
    for (int stg = 1; stg < 4 && rc == SHX_OK; ++stg) {
        rc = shx_system_rhs(s, w[stg - 1], &ws->k, t + dt * kRki[stg - 1]);
        if (rc != SHX_OK) break;
        field_lc(s, *u, ws->k, dt * kRki[stg], ws->w[stg - 1]);
        rc = shx_system_update_boundary(s, &ws->w[stg - 1], t + dt * kRki[stg]);
    }
    if (rc != SHX_OK) return rc;
    // backward: u' = BU(u + sum_s dt rkf[s] k_s), k_s = rhs(w_s)
    hipStream_t st = s->ctx->stream;
    auto copy = [&](const shx_field& a, shx_field& b) {
        hipMemcpyAsync(b.D, a.D, sizeof(double) * (size_t)s->N, hipMemcpyDeviceToDevice, st);
        for (int c = 0; c < 3; ++c)
            if (s->nR[c])
                hipMemcpyAsync(b.R[c], a.R[c], sizeof(double) * (size_t)s->nR[c], hipMemcpyDeviceToDevice, st);
    };
    copy(*ubar, ws->g);
    project(s, ws->g);  // g = P ubar'
    copy(ws->g, *ubar);  // ubar = g (u enters the final sum directly)
    field_scale(s, dt * kRkf[3], ws->g, ws->kb);  // kbar_3 = dt rkf[3] g
    // the coefficient gradient accumulates on the device through the four
    // stages (each stage's sum added to the running value, in stage order)
    if (coef_grad) {
        const hipError_t e = hipMemcpyAsync(ws->gacc, coef_grad, sizeof(double) * 27, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return hip_fail(s->ctx, e, "rk4 adjoint: coefficient gradient");
    }
    rc = check(s->ctx, "rk4 adjoint");
    for (int stg = 3; stg >= 0 && rc == SHX_OK; --stg) {
        rc = rhs_T(s, ws->kb, ws->wb, w[stg], coef_grad ? ws->gacc : nullptr);
        if (rc != SHX_OK) break;
        if (stg > 0) {
            project(s, ws->wb);
            field_add(s, ws->wb, *ubar);
            // kbar_{s-1} = dt rkf[s-1] g + dt rki[s] P wbar
            field_lc2(s, dt * kRkf[stg - 1], ws->g, dt * kRki[stg], ws->wb, ws->kb);
        } else {
            field_add(s, ws->wb, *ubar);
        }
        rc = check(s->ctx, "rk4 adjoint");
    }
    if (rc == SHX_OK && coef_grad) {
        const hipError_t e = hipMemcpyAsync(coef_grad, ws->gacc, sizeof(double) * 27, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(s->ctx, e, "rk4 adjoint: coefficient gradient");
    }
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_system_set_interior(shx_system* s, int dir, const double* coef, int width)
{
    if (!s || !coef || dir < 0 || dir > 2 || !s->d[dir]) return SHX_EINVAL;
    shx_derivative* d = s->d[dir];
    if (width != d->dev.width) return fail(s->ctx, SHX_EINVAL, "set_interior: width differs from the operator's");
    hipError_t e = hipMemcpyAsync(d->coef, coef, sizeof(double) * (size_t)width, hipMemcpyHostToDevice, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "set_interior");
    for (int m = 0; m < width; ++m) d->h_coef[m] = coef[m];
    destroy_transpose(d->T);  // its listed entries hold interior coefficients
    d->T = nullptr;
    return SHX_OK;
}

int shx_system_get_interior(const shx_system* s, int dir, double coef[9], int* width)
{
    if (!s || !coef || !width || dir < 0 || dir > 2 || !s->d[dir]) return SHX_EINVAL;
    const shx_derivative* d = s->d[dir];
    for (int m = 0; m < 9; ++m) coef[m] = m < d->dev.width ? d->h_coef[m] : 0.0;
    *width = d->dev.width;
    return SHX_OK;
}

}  // extern "C"
