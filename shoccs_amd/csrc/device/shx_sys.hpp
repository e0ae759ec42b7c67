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
    const unsigned* fluid_mask;
};

constexpr double kTwoPi = 6.283185307179586;  // 2 * pi, as 2 * std::numbers::pi

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


// Arguments of the x-marching fused stage kernels.
struct MarchArgs {
    DevSys s;
    DevOps ops;
    const double* us;
    const double* usR[3];
    const double* u0;
    double* acc;
    double* out;
    double c_acc, c_next, t_next;
    int nbz, nby, xchunk;
    unsigned nblocks;
};

// Launch one fused stage of the constant/radial scalar wave with the
// x-marching kernel (tile TY x 64, chunk of a.xchunk planes); returns -1 for
// an unsupported TY.
int launch_wave_stage_march(hipStream_t st, int stage, int ty, const MarchArgs& a);

}  // namespace shx
