// shx_mixed.hip -- the mixed-precision instantiation of the x-marching fused
// RK4 stage (shx_march.hpp): the stage register is FP32 and interior
// stencils on it are summed in FP32 with Kahan compensation; u0 and the RK
// accumulator stay FP64, as do closure rows, B terms (whose weights grow as
// the cut fraction goes to zero) and cut points.  Separate translation unit
// so that the two forms compile in parallel.
#include <hip/hip_runtime.h>

#include "shx_march.hpp"
#include "shx_sys.hpp"

namespace shx {

int launch_stage_march_mixed(hipStream_t st, int kind, int stage, int width, int ty, int mode,
                             const MarchArgs& a, const SpecialArgs& sp, const StageScalars& ss)
{
    return launch_stage_form<true>(st, kind, stage, width, ty, mode, a, sp, ss);
}

int launch_batch_stage_mixed(hipStream_t st, int kind, int stage, int width, int ty,
                             const BatchLaunch& bl)
{
    return launch_batch_form<true>(st, kind, stage, width, ty, bl);
}

}  // namespace shx
