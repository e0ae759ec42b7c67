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
    // host copies of the irregular columns' lists (the system merges the
    // three directions' lists into one)
    std::vector<int> h_irr_idx, h_irr_ptr, h_irr_row;
    std::vector<double> h_irr_w;
    // host copies of the closure rows' entries (column, row, weight): with the
    // circulant rows, whose columns every point can form by itself, they make
    // up O; the system merges the three directions' entries by column
    std::vector<int> h_cl_col, h_cl_row;
    std::vector<double> h_cl_w;
};

// Adjoint workspace of a system: stage inputs w1..w3, the stage-RHS scratch
// and the adjoint registers.
struct shx_adjoint_ws {
    shx_field w[3];
    shx_field k, g, kb, wb;
    double* part = nullptr;  // coefficient-gradient partials, 15 per workgroup, accumulated
    double* gacc = nullptr;  // coefficient gradient on the device, 3 x 9 (partials added on read)
    long npart = 0;
    int coef_w = 0;          // stencil width of the accumulated partials
    // the irregular columns of the three directions in one list: column,
    // entries (row, weight, direction), directions ascending then rows
    long nirr = 0;
    int* irr_col = nullptr;
    int* irr_ptr = nullptr;
    int* irr_row = nullptr;
    double* irr_w = nullptr;
    int8_t* irr_dir = nullptr;
    bool lists_ready = false;
    // per column, bit c W + m: row q + (W/2 - m) st_c is a circulant row of
    // direction c with a nonzero weight (built once; the mesh fixes it)
    unsigned short* tmask = nullptr;
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

// ---------------------------------------------- fused transposed rhs (all directions)
// The rhs weight of direction c at mesh point (i, j, k), by which the
// adjoint kbar is multiplied before that direction's transpose: heat k,
// scalar wave gG_c; zero on Dirichlet grid faces (heat.cpp:119-162,
// scalar_wave.cpp:208-238).
__device__ __forceinline__ double adj_weight(const DevSys& s, int c, int i, int j, int k)
{
    if (on_dirichlet_face(s, i, j, k)) return 0.0;
    if (s.kind == SHX_SYS_HEAT) return s.k;
    if (s.kind == SHX_SYS_WAVE_LINEAR) return -s.vel[c];
    return wave_neg_g(s, c, s.x[i], s.y[j], s.z[k]);
}

// Per direction of the fused transposes: the stored interior stencil, the
// stride, the extent and the mask of the points whose row is a circulant row.
struct AdjDir {
    const double* coef;
    long st;
    const unsigned* intr;
    int present;
};
struct AdjDirs {
    AdjDir d[3];
};

// Dirichlet grid face of axis a at coordinate q along it.
__device__ __forceinline__ bool face_of(const DevSys& s, int a, int q, int n)
{
    return (s.bc[2 * a] == SHX_BC_DIRICHLET && q == 0) || (s.bc[2 * a + 1] == SHX_BC_DIRICHLET && q == n - 1);
}

// Which rows of column q contribute to the circulant-row part of the
// transpose: bit c W + m is set when row p = q + (W/2 - m) st_c lies on the
// mesh, is a circulant row of direction c, and does not lie on a Dirichlet
// face (along its own axis, or along another axis, where the point is on one
// already): its weight is not zero.  Fixed by the mesh, so built once.
__global__ void k_tap_mask(DevSys s, AdjDirs ad, long N, int W, unsigned short* __restrict__ mask)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    const int hw = W / 2;
    int ijk[3];
    point_coords(q, s.n1, s.n2, ijk[0], ijk[1], ijk[2]);
    const int nn[3] = {s.n0, s.n1, s.n2};
    bool fa[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) fa[a] = face_of(s, a, ijk[a], nn[a]);
    unsigned m = 0u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const AdjDir& a = ad.d[c];
        if (!a.present) continue;
        const bool other = (c != 0 && fa[0]) || (c != 1 && fa[1]) || (c != 2 && fa[2]);
        for (int t = 0; t < W; ++t) {
            const int o = hw - t;
            const int qc = ijk[c] + o;
            if (qc < 0 || qc >= nn[c]) continue;
            if (!bit(a.intr, q + (long)o * a.st)) continue;
            if (other || face_of(s, c, qc, nn[c])) continue;
            m |= 1u << (c * W + t);
        }
    }
    mask[q] = (unsigned short)m;
}

// The circulant-row part of (d rhs / d u)^T kbar at every column q: for each
// direction c (in order) and tap m (in stencil order), the row p = q + (hw -
// m) st_c contributes coef_c[m] y_c[p] when it is a circulant row, with y_c =
// weight_c kbar (zero weight on Dirichlet faces).  The closure rows and the B
// rows follow from k_adj_lists.  With GRAD, the same products give the
// coefficient gradient: row p's term of coefficient m reads w at column q, so
// y_c[p] w[q] is added to value c W + m; per workgroup the values are summed
// in a fixed order and added to the workgroup's slot of `part` (one writer
// per slot and launch: deterministic).  The tap mask says which rows take
// part; the others add 0.0 (which leaves a sum started from 0.0 unchanged),
// reading column q itself so that every load stays in bounds.  CONSTW: the
// weight is constant per direction (heat: k, linear wave: -a_c); else the
// radial wave's weight at the row.
template <int W, bool GRAD, bool CONSTW>
__global__ __launch_bounds__(kT) void k_rhsT(DevSys s, AdjDirs ad, long N, const unsigned short* __restrict__ tmask,
                                             const double* __restrict__ kbar, double* __restrict__ wbar,
                                             const double* __restrict__ w, double* __restrict__ part)
{
    constexpr int hw = W / 2, NV = 3 * W;
    static_assert(NV * 16 <= kT, "one thread per partial sum");
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = q < N;
    double g[GRAD ? NV : 1];
    if constexpr (GRAD)
#pragma unroll
        for (int v = 0; v < NV; ++v) g[v] = 0.0;
    if (in) {
        const unsigned mk = tmask[q];
        const double wq = GRAD ? w[q] : 0.0;
        int ijk[3] = {0, 0, 0};
        if constexpr (!CONSTW) point_coords(q, s.n1, s.n2, ijk[0], ijk[1], ijk[2]);
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const AdjDir& a = ad.d[c];
            if (!a.present) continue;
            const double wc = s.kind == SHX_SYS_HEAT ? s.k : -s.vel[c];
            double sc = 0.0;
#pragma unroll
            for (int m = 0; m < W; ++m) {
                const int o = hw - m;
                const bool on = (mk >> (c * W + m)) & 1u;
                const long p = on ? q + (long)o * a.st : q;
                double wgt = wc;
                if constexpr (!CONSTW)
                    wgt = on ? adj_weight(s, c, ijk[0] + (c == 0 ? o : 0), ijk[1] + (c == 1 ? o : 0),
                                          ijk[2] + (c == 2 ? o : 0))
                             : 0.0;
                const double kb = kbar[p];
                const double y = on ? wgt * kb : 0.0;
                sc += a.coef[m] * y;
                if constexpr (GRAD) g[c * W + m] += y * wq;
            }
            acc += sc;
        }
        wbar[q] = acc;
    }
    if constexpr (GRAD) {
        // the workgroup's sums in a fixed order: lane values staged in LDS,
        // 16 partial sums of 16 consecutive lanes per value, then those 16
        __shared__ double red[NV][kT];
        __shared__ double red2[NV][16];
#pragma unroll
        for (int v = 0; v < NV; ++v) red[v][threadIdx.x] = g[v];
        __syncthreads();
        if ((int)threadIdx.x < NV * 16) {
            const int v = threadIdx.x / 16, q0 = (threadIdx.x % 16) * 16;
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < 16; ++r) sum += red[v][q0 + r];
            red2[v][threadIdx.x % 16] = sum;
        }
        __syncthreads();
        if ((int)threadIdx.x < NV) {
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < 16; ++r) sum += red2[threadIdx.x][r];
            part[(long)blockIdx.x * NV + threadIdx.x] += sum;
        }
    }
}

// The listed rows in one launch: items [0, ncol) are the columns read by
// closure rows (wbar.D[col] += their entries' w y_dir[row], directions
// ascending), then the cut points of R(0), R(1), R(2) (wbar.R[c][e] = B^T
// entries' sum).
struct AdjLists {
    long ncol;
    const int* col;
    const int* ptr;
    const int* row;
    const double* w;
    const int8_t* dir;
    long off[4];  // cut points of R(c): items ncol + [off[c], off[c + 1])
    const int* bptr[3];
    const int* brow[3];
    const double* bw[3];
    double* xR[3];
};
__global__ void k_adj_lists(DevSys s, AdjLists L, const double* __restrict__ kbar, double* __restrict__ wbar)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= L.ncol + L.off[3]) return;
    double sum = 0.0;
    if (g < L.ncol) {
        for (int t = L.ptr[g]; t < L.ptr[g + 1]; ++t) {
            const long r = L.row[t];
            int i, j, k;
            point_coords(r, s.n1, s.n2, i, j, k);
            sum += L.w[t] * (adj_weight(s, L.dir[t], i, j, k) * kbar[r]);
        }
        wbar[L.col[g]] += sum;
        return;
    }
    const long gr = g - L.ncol;
    const int c = gr < L.off[1] ? 0 : gr < L.off[2] ? 1 : 2;
    const long e = gr - L.off[c];
    const int* ptr = c == 0 ? L.bptr[0] : c == 1 ? L.bptr[1] : L.bptr[2];
    const int* row = c == 0 ? L.brow[0] : c == 1 ? L.brow[1] : L.brow[2];
    const double* w = c == 0 ? L.bw[0] : c == 1 ? L.bw[1] : L.bw[2];
    double* x = c == 0 ? L.xR[0] : c == 1 ? L.xR[1] : L.xR[2];
    if (ptr)
        for (int t = ptr[e]; t < ptr[e + 1]; ++t) {
            const long r = row[t];
            int i, j, k;
            point_coords(r, s.n1, s.n2, i, j, k);
            sum += w[t] * (adj_weight(s, c, i, j, k) * kbar[r]);
        }
    x[e] = sum;
}

// gacc[9 c + m] += sum over the nb workgroup slots of value c W + m (fixed
// order: lane l sums slots l, l + 64, ..., then a butterfly), and the slots
// are cleared; one wave per value.
__global__ void k_coef_reduce3(int nb, int W, double* __restrict__ part, double* __restrict__ gacc)
{
    const int v = blockIdx.x, lane = threadIdx.x;
    double sum = 0.0;
    for (int b = lane; b < nb; b += 64) {
        sum += part[(long)b * 3 * W + v];
        part[(long)b * 3 * W + v] = 0.0;
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (lane == 0) gacc[9 * (v / W) + v % W] += sum;
}

// A field's D and R parts as one index space of N + nR[0] + nR[1] + nR[2]
// items (the R sets in order).
struct FieldParts {
    long N;
    long off[4];  // R(c): items N + [off[c], off[c + 1])
    const signed char* dirich[3];
};
struct PartPtr {
    double* D;
    double* R[3];
};
// item g of the parts: the element pointer, and whether P zeroes it
// (Dirichlet grid face of D, Dirichlet cut point)
__device__ __forceinline__ double* part_at(const DevSys& s, const FieldParts& f, const PartPtr& p, long g, bool& zero)
{
    if (g < f.N) {
        int i, j, k;
        point_coords(g, s.n1, s.n2, i, j, k);
        zero = on_dirichlet_face(s, i, j, k);
        return p.D + g;
    }
    const long gr = g - f.N;
    const int c = gr < f.off[1] ? 0 : gr < f.off[2] ? 1 : 2;
    const long e = gr - f.off[c];
    zero = (c == 0 ? f.dirich[0] : c == 1 ? f.dirich[1] : f.dirich[2])[e] != 0;
    return (c == 0 ? p.R[0] : c == 1 ? p.R[1] : p.R[2]) + e;
}

// End of a backward stage: pw = P wbar, ubar += pw, kb = ca g + cb pw (P is
// the identity when proj is false: the last stage, whose kb is not used).
__global__ void k_adj_stage_end(DevSys s, FieldParts f, PartPtr wbar, PartPtr g, PartPtr ubar, PartPtr kb, double ca,
                                double cb, int proj)
{
    const long gi = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= f.N + f.off[3]) return;
    bool zero;
    const double* wp = part_at(s, f, wbar, gi, zero);
    const double pw = proj && zero ? 0.0 : *wp;
    *part_at(s, f, ubar, gi, zero) += pw;
    if (proj) *part_at(s, f, kb, gi, zero) = ca * *part_at(s, f, g, gi, zero) + cb * pw;
}

// Start of a step's backward sweep: g = P ubar, ubar = g, kb = c g.
__global__ void k_adj_step_begin(DevSys s, FieldParts f, PartPtr ubar, PartPtr g, PartPtr kb, double c)
{
    const long gi = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= f.N + f.off[3]) return;
    bool zero;
    double* up = part_at(s, f, ubar, gi, zero);
    const double v = zero ? 0.0 : *up;
    *up = v;
    *part_at(s, f, g, gi, zero) = v;
    *part_at(s, f, kb, gi, zero) = c * v;
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
    std::vector<Ent> ent, bent, cl;
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
                for (int jj = 0; jj < ncols; ++jj) {
                    ent.push_back({base + (long)(col0 + jj) * st, p, a[jj]});
                    cl.push_back({base + (long)(col0 + jj) * st, p, a[jj]});
                }
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
    for (const Ent& x : cl) {
        T->h_cl_col.push_back((int)x.col);
        T->h_cl_row.push_back((int)x.row);
        T->h_cl_w.push_back(x.w);
    }
    T->h_irr_idx = irr_idx;
    T->h_irr_ptr = irr_ptr;
    T->h_irr_row = irr_row;
    T->h_irr_w = irr_w;
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
    if (rc == SHX_OK) {
        w->npart = std::max<long>(1, (s->N + kT - 1) / kT);  // one slot per workgroup of k_rhsT
        hipError_t e = hipMalloc((void**)&w->part, sizeof(double) * (size_t)w->npart * 27);
        if (e == hipSuccess) e = hipMalloc((void**)&w->gacc, sizeof(double) * 27);
        if (e == hipSuccess) e = hipMemset(w->part, 0, sizeof(double) * (size_t)w->npart * 27);
        if (e == hipSuccess) e = hipMemset(w->gacc, 0, sizeof(double) * 27);
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

// The closure-row entries of the system's directions merged by column
// (built once, after the directions' transposes).
int ensure_lists(shx_system* s)
{
    shx_adjoint_ws* ws = s->adj;
    if (ws->lists_ready) return SHX_OK;
    struct Ent {
        int col, c, row;
        double w;
    };
    std::vector<Ent> ent;
    for (int c = 0; c < 3; ++c) {
        shx_derivative* d = s->d[c];
        if (!d || !d->dev.present || !d->T) continue;
        const shx_deriv_T* T = d->T;
        for (size_t e = 0; e < T->h_cl_col.size(); ++e) ent.push_back({T->h_cl_col[e], c, T->h_cl_row[e], T->h_cl_w[e]});
    }
    // by column, then direction; rows ascending within (stable: the entries
    // of a direction come row by row)
    std::stable_sort(ent.begin(), ent.end(),
                     [](const Ent& a, const Ent& b) { return a.col != b.col ? a.col < b.col : a.c < b.c; });
    std::vector<int> col, ptr{0}, row;
    std::vector<double> w;
    std::vector<int8_t> dir;
    for (size_t e = 0; e < ent.size(); ++e) {
        if (e == 0 || ent[e].col != ent[e - 1].col) {
            if (e) ptr.push_back((int)row.size());
            col.push_back(ent[e].col);
        }
        row.push_back(ent[e].row);
        w.push_back(ent[e].w);
        dir.push_back((int8_t)ent[e].c);
    }
    if (!ent.empty()) ptr.push_back((int)row.size());
    ws->nirr = (long)col.size();
    int rc = up(s->ctx, &ws->irr_col, col);
    if (rc == SHX_OK) rc = up(s->ctx, &ws->irr_ptr, ptr);
    if (rc == SHX_OK) rc = up(s->ctx, &ws->irr_row, row);
    if (rc == SHX_OK) rc = up(s->ctx, &ws->irr_w, w);
    if (rc == SHX_OK) rc = up(s->ctx, &ws->irr_dir, dir);
    ws->lists_ready = rc == SHX_OK;
    return rc;
}

// wbar = (d rhs / d u)^T kbar, and, when grad is set, the stage's
// coefficient-gradient terms with stage input w accumulated in the partials:
// one pass over the columns (circulant rows of every direction, and the
// gradient), one over the listed rows (closure columns, B^T).
int rhs_T(shx_system* s, const shx_field& kbar, shx_field& wbar, const shx_field* w, bool grad)
{
    shx_adjoint_ws* ws = s->adj;
    hipStream_t st = s->ctx->stream;
    AdjDirs ad{};
    int W = 0;
    int rc = SHX_OK;
    for (int c = 0; c < 3 && rc == SHX_OK; ++c) {
        shx_derivative* d = s->d[c];
        if (!d || !d->dev.present) continue;
        rc = build_transpose(d);
        if (rc != SHX_OK) break;
        if (W && d->dev.width != W) return fail(s->ctx, SHX_EUNSUPPORTED, "adjoint: directions differ in stencil width");
        W = d->dev.width;
        ad.d[c] = AdjDir{d->coef, d->dev.stride, d->T->int_mask, 1};
    }
    if (rc != SHX_OK) return rc;
    for (int c = 0; c < 3; ++c)
        if (s->nR[c] && !ad.d[c].present) k_fill0<<<blocks_for(s->nR[c]), kT, 0, st>>>(s->nR[c], wbar.R[c]);
    if (W == 0) {  // no direction: the rhs does not depend on u
        k_fill0<<<blocks_for(s->N), kT, 0, st>>>(s->N, wbar.D);
        return check(s->ctx, "adjoint");
    }
    if (grad && ws->coef_w && ws->coef_w != W)
        return fail(s->ctx, SHX_EUNSUPPORTED, "adjoint: stencil width changed");
    grad = grad && w;
    if (grad) ws->coef_w = W;
    const unsigned nb = blocks_for(s->N);
    if (W != 3 && W != 5) return fail(s->ctx, SHX_EUNSUPPORTED, "adjoint: interior width must be 3 or 5");
    if (!ws->tmask) {
        hipError_t e = hipMalloc((void**)&ws->tmask, sizeof(unsigned short) * (size_t)std::max<long>(s->N, 1));
        if (e != hipSuccess) return hip_fail(s->ctx, e, "adjoint tap mask");
        k_tap_mask<<<nb, kT, 0, st>>>(s->dev, ad, s->N, W, ws->tmask);
    }
    const bool constw = s->kind == SHX_SYS_HEAT || s->kind == SHX_SYS_WAVE_LINEAR;
#define SHX_RHST(WW, CW)                                                                                  \
    if (grad) k_rhsT<WW, true, CW><<<nb, kT, 0, st>>>(s->dev, ad, s->N, ws->tmask, kbar.D, wbar.D, w->D, ws->part); \
    else k_rhsT<WW, false, CW><<<nb, kT, 0, st>>>(s->dev, ad, s->N, ws->tmask, kbar.D, wbar.D, nullptr, nullptr);
    if (W == 3) {
        if (constw) { SHX_RHST(3, true) }
        else { SHX_RHST(3, false) }
    } else {
        if (constw) { SHX_RHST(5, true) }
        else { SHX_RHST(5, false) }
    }
#undef SHX_RHST
    rc = ensure_lists(s);
    if (rc != SHX_OK) return rc;
    AdjLists L{};
    L.ncol = ws->nirr;
    L.col = ws->irr_col;
    L.ptr = ws->irr_ptr;
    L.row = ws->irr_row;
    L.w = ws->irr_w;
    L.dir = ws->irr_dir;
    for (int c = 0; c < 3; ++c) {
        // cut points of an absent direction were zero-filled above
        L.off[c + 1] = L.off[c] + (ad.d[c].present ? s->nR[c] : 0);
        if (ad.d[c].present) {
            const shx_deriv_T* T = s->d[c]->T;
            L.bptr[c] = T->b_ptr;
            L.brow[c] = T->b_row;
            L.bw[c] = T->b_w;
        }
        L.xR[c] = ad.d[c].present ? wbar.R[c] : nullptr;
    }
    const long items = L.ncol + L.off[3];
    if (items > 0) k_adj_lists<<<blocks_for(items), kT, 0, st>>>(s->dev, L, kbar.D, wbar.D);
    return check(s->ctx, "transposed rhs");
}

// Adds the accumulated coefficient partials to gacc and clears them.
int coef_flush(shx_system* s)
{
    shx_adjoint_ws* ws = s->adj;
    if (ws->coef_w == 0) return SHX_OK;
    k_coef_reduce3<<<3 * ws->coef_w, 64, 0, s->ctx->stream>>>((int)ws->npart, ws->coef_w, ws->part, ws->gacc);
    return check(s->ctx, "adjoint: coefficient gradient");
}

// The backward sweep of one RK4 step over its stage inputs w[0..3] (w[0] the
// step's input): on entry ubar = dJ/du_{n+1}, on exit dJ/du_n; the
// coefficient gradient (when grad is set) accumulates on the device.
int rk4_backward(shx_system* s, const shx_field* const w[4], shx_field& ubar, double dt, bool grad);

constexpr double kRki[4] = {0.0, 0.5, 0.5, 1.0};
constexpr double kRkf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

int rk4_backward(shx_system* s, const shx_field* const w[4], shx_field& ubar, double dt, bool grad)
{
    shx_adjoint_ws* ws = s->adj;
    hipStream_t st = s->ctx->stream;
    FieldParts f{};
    f.N = s->N;
    for (int r = 0; r < 3; ++r) {
        f.off[r + 1] = f.off[r] + s->nR[r];
        f.dirich[r] = s->rdir[r];
    }
    auto pp = [](const shx_field& x) {
        PartPtr p{x.D, {x.R[0], x.R[1], x.R[2]}};
        return p;
    };
    const unsigned nb = blocks_for(f.N + f.off[3]);
    // u' = BU(u + sum_s dt rkf[s] k_s): g = P ubar', ubar = g (u enters the
    // final sum directly), kbar_3 = dt rkf[3] g
    k_adj_step_begin<<<nb, kT, 0, st>>>(s->dev, f, pp(ubar), pp(ws->g), pp(ws->kb), dt * kRkf[3]);
    int rc = check(s->ctx, "rk4 adjoint");
    for (int stg = 3; stg >= 0 && rc == SHX_OK; --stg) {
        rc = rhs_T(s, ws->kb, ws->wb, w[stg], grad);
        if (rc != SHX_OK) break;
        // stg > 0: ubar += P wbar, kbar_{s-1} = dt rkf[s-1] g + dt rki[s] P wbar;
        // stg = 0: ubar += wbar (kb is not needed after the last stage)
        const double ca = stg > 0 ? dt * kRkf[stg - 1] : 0.0, cb = stg > 0 ? dt * kRki[stg] : 0.0;
        k_adj_stage_end<<<nb, kT, 0, st>>>(s->dev, f, pp(ws->wb), pp(ws->g), pp(ubar), pp(ws->kb), ca, cb, stg > 0);
        rc = check(s->ctx, "rk4 adjoint");
    }
    return rc;
}

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
    hipFree(w->part);
    hipFree(w->gacc);
    hipFree(w->irr_col);
    hipFree(w->irr_ptr);
    hipFree(w->irr_row);
    hipFree(w->irr_w);
    hipFree(w->irr_dir);
    hipFree(w->tmask);
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
    if (rc == SHX_OK) rc = rhs_T(s, *kbar, *wbar, nullptr, false);
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
    // the coefficient gradient accumulates on the device through the four
    // stages (each stage's sum added to the running value, in stage order)
    hipStream_t st = s->ctx->stream;
    if (coef_grad) {
        const hipError_t e = hipMemcpyAsync(ws->gacc, coef_grad, sizeof(double) * 27, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return hip_fail(s->ctx, e, "rk4 adjoint: coefficient gradient");
    }
    rc = rk4_backward(s, w, *ubar, dt, coef_grad != nullptr);
    if (rc == SHX_OK && coef_grad) rc = coef_flush(s);
    if (rc == SHX_OK && coef_grad) {
        const hipError_t e = hipMemcpyAsync(coef_grad, ws->gacc, sizeof(double) * 27, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return hip_fail(s->ctx, e, "rk4 adjoint: coefficient gradient");
    }
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_rk4_adjoint_stages_async(shx_system* s, const shx_field* w, shx_field* ubar, double dt, int with_coef)
{
    if (!s || !w || !ubar) return SHX_EINVAL;
    if (s->general || s->hm)
        return fail(s->ctx, SHX_EUNSUPPORTED,
                    "adjoint: floating/Neumann bodies, Neumann grid faces and host-evaluated solutions are not supported");
    int rc = ensure_ws(s);
    if (rc != SHX_OK) return rc;
    const shx_field* ws[4] = {&w[0], &w[1], &w[2], &w[3]};
    return rk4_backward(s, ws, *ubar, dt, with_coef != 0);
}

int shx_adjoint_prepare(shx_system* s)
{
    if (!s) return SHX_EINVAL;
    if (s->general || s->hm)
        return fail(s->ctx, SHX_EUNSUPPORTED,
                    "adjoint: floating/Neumann bodies, Neumann grid faces and host-evaluated solutions are not supported");
    int rc = ensure_ws(s);
    for (int c = 0; c < 3 && rc == SHX_OK; ++c)
        if (s->d[c] && s->d[c]->dev.present) rc = build_transpose(s->d[c]);
    if (rc == SHX_OK) rc = ensure_lists(s);
    return rc == SHX_OK ? sync(s->ctx) : rc;
}

int shx_adjoint_coef_set(shx_system* s, const double* g)
{
    if (!s || !g) return SHX_EINVAL;
    int rc = ensure_ws(s);
    if (rc != SHX_OK) return rc;
    rc = coef_flush(s);  // partials not yet added are dropped by the overwrite
    if (rc != SHX_OK) return rc;
    const hipError_t e = hipMemcpyAsync(s->adj->gacc, g, sizeof(double) * 27, hipMemcpyHostToDevice, s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "adjoint coefficient gradient");
    return sync(s->ctx);
}

int shx_adjoint_coef_get(shx_system* s, double* g)
{
    if (!s || !g) return SHX_EINVAL;
    int rc = ensure_ws(s);
    if (rc != SHX_OK) return rc;
    rc = coef_flush(s);
    if (rc != SHX_OK) return rc;
    const hipError_t e = hipMemcpyAsync(g, s->adj->gacc, sizeof(double) * 27, hipMemcpyDeviceToHost, s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "adjoint coefficient gradient");
    return sync(s->ctx);
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
