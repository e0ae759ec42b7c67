// shx_march.hip -- the x-marching form of one fused RK4 stage: the per-point
// classification kernel, and the FP64 form's instantiation of the tile and
// special-point templates of shx_march.hpp.
//
// Every point of D carries a class byte, computed once per system (interior,
// no rows, closure pattern or special; see shx_sys.hpp).  The FP64 form uses
// the formulas of shx_point.hpp in the reference's order, so its results are
// bit-identical to the point-wise tiles.
#include <hip/hip_runtime.h>

#include "shx_march.hpp"
#include "shx_point.hpp"
#include "shx_sys.hpp"

namespace shx {

namespace {

__device__ __forceinline__ bool in_fast(unsigned long long fd, int q)
{
    const int lo0 = (int)(fd & 0xffffu), hi0 = (int)((fd >> 16) & 0xffffu);
    const int lo1 = (int)((fd >> 32) & 0xffffu), hi1 = (int)(fd >> 48);
    return (q >= lo0 && q < hi0) || (q >= lo1 && q < hi1);
}

// Whether line `pair` of an operator has a row at position q.
__device__ __forceinline__ bool row_exists(const DevDeriv& d, int pair, int q)
{
    const int b_end = d.pair_ptr[pair + 1];
    for (int b = d.pair_ptr[pair]; b < b_end; ++b) {
        const shx_line_block& blk = d.blocks[b];
        const int lr = q - blk.lo;
        if (lr < 0) return false;
        if (lr < blk.nl + blk.ni + blk.nr) return true;
    }
    return false;
}

// 0 = no row, 1 = interior (circulant) row, 2 = other row
__device__ __forceinline__ int row_kind(const DevDeriv& d, int pair, int q)
{
    if (in_fast(d.fast[pair], q)) return 1;
    return row_exists(d, pair, q) ? 2 : 0;
}

__global__ void k_classify(DevSys s, DevOps ops, unsigned char* __restrict__ cls, long N)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    int i, j, k;
    point_coords(idx, s.n1, s.n2, i, j, k);
    int present = 0, interior = 0, none = 0;
    if (ops.d[0].present) {
        ++present;
        const int r = row_kind(ops.d[0], j * s.n2 + k, i);
        interior += r == 1;
        none += r == 0;
    }
    if (ops.d[1].present) {
        ++present;
        const int r = row_kind(ops.d[1], i * s.n2 + k, j);
        interior += r == 1;
        none += r == 0;
    }
    if (ops.d[2].present) {
        ++present;
        const int r = row_kind(ops.d[2], i * s.n1 + j, k);
        interior += r == 1;
        none += r == 0;
    }
    const bool face = on_dirichlet_face(s, i, j, k);
    const bool fluid = is_fluid(s, idx);
    unsigned char c = kRawSpecial | (face ? kRawFace : 0) | (fluid ? kRawFluid : 0);
    if (!face && fluid && interior == present) c = kInterior;
    else if (!face && !fluid && none == present) c = kNoRows;
    cls[idx] = c;
}

__global__ void k_batch_advance(BatchClock* __restrict__ clk, int nprob)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nprob) clk[p].n += 1;
}

}  // namespace

int launch_classify(hipStream_t st, const DevSys& s, const DevOps& ops, unsigned char* cls, long N)
{
    if (N <= 0) return 0;
    k_classify<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(s, ops, cls, N);
    return 0;
}

int launch_stage_march(hipStream_t st, int kind, int stage, int width, int ty, int mode,
                       const MarchArgs& a, const SpecialArgs& sp, const StageScalars& ss)
{
    return launch_stage_form<false>(st, kind, stage, width, ty, mode, a, sp, ss);
}

int launch_batch_stage(hipStream_t st, int kind, int stage, int width, int ty, const BatchLaunch& bl)
{
    return launch_batch_form<false>(st, kind, stage, width, ty, bl);
}

int launch_batch_advance(hipStream_t st, BatchClock* clk, int nprob)
{
    if (nprob <= 0) return 0;
    k_batch_advance<<<(unsigned)((nprob + 255) / 256), 256, 0, st>>>(clk, nprob);
    return 0;
}

}  // namespace shx
