// shx_sys.hpp -- device view of a system and the point-wise pieces shared by
// the fused right-hand-side and RK4 stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "shx_internal.hpp"

namespace shx {

// Device view of a system (kernel argument).
struct DevSys {
    int kind;
    int n0, n1, n2;
    const double* x;
    const double* y;
    const double* z;
    int bc[6];
    double center[3], radius;
    double vel[3], x0[3], width, amp;
    double k;
    int mms_kind, mms_dims, mms_n;
    const double* mms;
    // sin-product MMS: sin(2 pi x_i), sin(2 pi y_j), sin(2 pi z_k) per mesh
    // coordinate (computed on the device with the device's sin)
    const double* sinx;
    const double* siny;
    const double* sinz;
    const unsigned* fluid_mask;
};

constexpr double kTwoPi = 6.283185307179586;  // 2 * pi, as 2 * std::numbers::pi

// Kernel family of a system: SHX_SYS_* for the public kinds, with heat split
// by manufactured solution (the sin product reads per-axis tables; the
// gaussian is evaluated per point out of line).
constexpr int kHeatGauss = 3;
template <int KIND>
constexpr bool kIsHeat = KIND == SHX_SYS_HEAT || KIND == kHeatGauss;

__device__ __forceinline__ double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

// scalar_wave.cpp:35-40 (radial) and the translated gaussian of the
// constant-velocity wave.
__device__ __forceinline__ double wave_exact(const DevSys& s, double x, double y, double z, double t)
{
    if (s.kind == SHX_SYS_WAVE_RADIAL) {
        const double r[3] = {x - s.center[0], y - s.center[1], z - s.center[2]};
        return sin(kTwoPi * (sqrt(dot3(r, r)) - s.radius - t));
    }
    const double q[3] = {x - s.x0[0] - s.vel[0] * t, y - s.x0[1] - s.vel[1] * t,
                         z - s.x0[2] - s.vel[2] * t};
    return s.amp * exp(-dot3(q, q) / (2 * s.width * s.width));
}

// gG component (scalar_wave.cpp:28-33): -(x - c)/|x - c| or -a.
__device__ __forceinline__ double wave_neg_g(const DevSys& s, int comp, double x, double y, double z)
{
    if (s.kind == SHX_SYS_WAVE_RADIAL) {
        const double r[3] = {x - s.center[0], y - s.center[1], z - s.center[2]};
        const double lc = comp == 0 ? x : comp == 1 ? y : z;
        return -(lc - s.center[comp]) / sqrt(dot3(r, r));
    }
    return -s.vel[comp];
}

// point on a Dirichlet grid face (for_each_grid_bc_desc, selection_desc.hpp:211-222)
__device__ __forceinline__ bool on_dirichlet_face(const DevSys& s, int i, int j, int k)
{
    return (s.bc[0] == SHX_BC_DIRICHLET && i == 0) || (s.bc[1] == SHX_BC_DIRICHLET && i == s.n0 - 1) ||
           (s.bc[2] == SHX_BC_DIRICHLET && j == 0) || (s.bc[3] == SHX_BC_DIRICHLET && j == s.n1 - 1) ||
           (s.bc[4] == SHX_BC_DIRICHLET && k == 0) || (s.bc[5] == SHX_BC_DIRICHLET && k == s.n2 - 1);
}

__device__ __forceinline__ bool is_fluid(const DevSys& s, long idx)
{
    return (s.fluid_mask[idx >> 5] >> (idx & 31)) & 1u;
}

struct DevOps {
    DevDeriv d[3];
};

struct RPtrs {
    const double* r[3];
};


// Per-point class of the x-marching fused form (one byte per point of D):
//   kInterior -- every present direction has an interior circulant row here,
//                not on a Dirichlet face, fluid: register/LDS stencils only;
//   kNoRows   -- no row in any direction, not on a face, not fluid (solid
//                points): the right-hand side of zero derivatives;
//   kPattern0 + (ix (ny + 1) + iy) (nz + 1) + iz -- fluid, not on a face,
//                each direction's row interior (index 0) or closure row
//                i - 1 of that direction's table: closure rows without B
//                terms whose columns the tile holds (x: the register queue,
//                y/z: the LDS plane);
//   kSpecial  -- anything else (rows with B terms, closures out of reach,
//                Dirichlet faces): written by the special-point work, from
//                a record, or for a face point (whose right-hand side is
//                zero) as exact(t_next) from its face item.
// The classification kernel marks raw special points kRawSpecial | flags;
// the host turns them into a pattern or kSpecial.
constexpr unsigned char kInterior = 0;
constexpr unsigned char kSpecial = 1;
constexpr unsigned char kNoRows = 2;
constexpr unsigned char kPattern0 = 3;
constexpr int kMaxPatterns = 256 - kPattern0;
constexpr unsigned char kRawSpecial = 0x80;
constexpr unsigned char kRawFace = 0x01;
constexpr unsigned char kRawFluid = 0x02;
constexpr int kMaxClosureRows = 16;   // per direction, staged in LDS
constexpr int kMaxClosureCols = 8;
// The register queue of the tiles holds hw + kXReach planes either side of
// the point; x closure rows whose columns lie outside it (the wall planes)
// are read from the stage input at their own planes, in the tiles.  A reach
// of 2 measured no faster at 256^3 than 0 and costs registers.
constexpr int kXReach = 0;

// A closure row without B terms: the dot of w[c] with the values at offsets
// off + c along its direction, c < n (w copied from the coefficient pool).
struct ClosureRow {
    int off, n;
    double w[kMaxClosureCols];
};

// Arguments of the x-marching fused stage kernel.  The stage input `us` and
// output `out` are FP64, or FP32 for the mixed form's stage register.
struct MarchArgs {
    DevSys s;
    DevOps ops;
    const void* us;
    const double* u0;
    double* acc;
    void* out;
    const unsigned char* cls;
    const ClosureRow* rows;  // 3 x kMaxClosureRows
    int nrows[3];
    int fixw[3];  // every table row of the direction has as many columns as the interior rows
    int nbz, nby, xchunk;
    int vec;  // points per lane along z: 2 (even z extent) or 1
    unsigned nblocks;
};

// A row of one direction's operator at a special point, other than an
// interior circulant row: dot of pool[coef + c] with the D values at
// offsets (off + c) * stride from the point, c < ncols, then nb (0-2) B
// terms pool[bw] * R[bobj] in the order the reference adds them.
struct SpecialRow {
    int off, ncols, coef, nb;
    int bobj0, bw0, bobj1, bw1;
};
constexpr unsigned kRowNone = 0xffffffffu;      // no row in this direction
constexpr unsigned kRowInterior = 0xfffffffeu;  // interior circulant row

// One special point: its index in D and, per direction, kRowNone,
// kRowInterior or an index into that direction's SpecialRow table.
struct SpecialRec {
    unsigned idx;
    unsigned ref[3];
};

// One free cut point (a cut point of a non-Dirichlet body) of the fused
// form: point q of R(set), and the entries of its rows of the three
// derivatives (n[0] of the x row, then n[1] of the y row, n[2] of the z row;
// a missing row has none).  Free cut point f's first kFreeSlots entries sit
// at slots f kFreeSlots + e of the entry list (unused slots name D point 0),
// so that they load together with its record; the rest continue from
// `over`.
constexpr int kFreeSlots = 12;
struct FreeRec {
    long q;
    long over;
    int n[3];
    int set;
};

// One entry of a free cut point's row (the derivative's R-space row entry
// in stored order): column >= 0 names a point of D, -1 - c cut point c of
// the row's set.
struct FreeEnt {
    long col;
    double w;
};

// Arguments of the special-point work: the free cut points (their rows, the
// rhs and the stage update), the listed points of D, the Dirichlet grid
// faces (exact(t_next)), then the cut points of the three directions
// (Dirichlet ones take exact(t_next)).
struct SpecialArgs {
    DevSys s;
    DevOps ops;
    const void* us;
    RPtrs usR;
    const double* u0;
    double* acc;
    void* out;
    const SpecialRec* recs;
    const SpecialRow *rows0, *rows1, *rows2;
    long nspecial;
    long nR0, nR1, nR2;
    const double *pos0, *pos1, *pos2;
    const signed char *dir0, *dir1, *dir2;
    double *outR0, *outR1, *outR2;
    long nfree;
    const FreeRec* frees;
    const FreeEnt* fents;
    long nface;  // face items: the sizes of the Dirichlet faces, summed
    RPtrs u0R;      // the step's u0 at the cut points
    double* accF;   // accumulator of the free cut points (nfree values)
};

// Per-stage scalars of one fused RK4 stage (rk4.cpp:27-46): the stage time,
// the accumulator and next-stage coefficients, and the time of the values
// the stage writes.
struct StageScalars {
    double t_s, c_acc, c_next, t_next;
};

// One problem of a batch, for one stage: its tile and special-point
// arguments and its workgroup groups (mode as launch_stage_march's).
struct BatchEntry {
    MarchArgs m;
    SpecialArgs sp;
    unsigned sg, mg;
    int mode;
};

// Time of one problem of a batch: t = t0 + n dt at the start of a step.
struct BatchClock {
    double t0, dt;
    long long n;
};

// One batched stage launch: per-problem entries of that stage, clocks, and
// the first workgroup of each problem (nprob entries; grid in total).
struct BatchLaunch {
    const BatchEntry* ent;
    const BatchClock* clk;
    const unsigned* start;
    int nprob;
    int vec;  // points per lane along z of every problem's tiles
    unsigned grid;
};

// Host launchers (shx_march.hip); each returns 0, or -1 for a template
// combination that is not built.  launch_stage_march runs one fused RK4 stage
// of the x-marching form: tiles and special items in one launch (mode 1:
// special workgroups first, 2: last, 3: spread evenly) or in two (mode 0).
int launch_classify(hipStream_t st, const DevSys& s, const DevOps& ops, unsigned char* cls, long N);
int launch_stage_march(hipStream_t st, int kind, int stage, int width, int ty, int mode,
                       const MarchArgs& a, const SpecialArgs& sp, const StageScalars& ss);
// The mixed form (shx_mixed.hip): stage register FP32, interior stencils on it
// in FP32 with compensated sums; u0, the accumulator, closure rows, B terms
// and cut points in FP64.
int launch_stage_march_mixed(hipStream_t st, int kind, int stage, int width, int ty, int mode,
                             const MarchArgs& a, const SpecialArgs& sp, const StageScalars& ss);
// One fused RK4 stage of every problem of a batch in one launch, and the
// clock advance after a step (n += 1 for every problem).
int launch_batch_stage(hipStream_t st, int kind, int stage, int width, int ty, const BatchLaunch& bl);
int launch_batch_stage_mixed(hipStream_t st, int kind, int stage, int width, int ty,
                             const BatchLaunch& bl);
int launch_batch_advance(hipStream_t st, BatchClock* clk, int nprob);

}  // namespace shx
