// shx_system_state.hpp -- host-side state behind the opaque shx_system handle
// (shared by the translation units that implement the system entry points).
#pragma once

#include <vector>

#include "shx_internal.hpp"
#include "shx_state.hpp"
#include "shx_sys.hpp"

struct shx_adjoint_ws;

// Heat with a manufactured solution whose functions are called on the host
// (SHX_MMS_HOST, the reference's lua_mms): host copies of the evaluation
// sites, and staging for the values copied to the device.
struct HostMms {
    shx_mms_fns f{};
    std::vector<double> xyz[3];          // mesh coordinates per axis
    std::vector<double> pos[3];          // cut-point positions, 3 per point
    std::vector<signed char> dirich[3];  // cut point of a Dirichlet body
    std::vector<int> fluid, faces, neu;  // fluid gather, Dirichlet faces, Neumann faces
    long neu_off[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<double> h;               // host staging
    double* d = nullptr;                 // device staging (max(N, nR) values)
    double* exD = nullptr;               // exact values on the fluid gather (statistics)
};

struct shx_system {
    shx_ctx* ctx = nullptr;
    int kind = 0;
    int family = 0;  // kernel family: kind, or kHeatGauss for heat with the gaussian MMS
    int n[3] = {1, 1, 1};
    long N = 0;
    long nR[3] = {0, 0, 0};
    shx_derivative* d[3] = {nullptr, nullptr, nullptr};
    shx::DevSys dev;
    shx::DevOps ops;
    shx::Tiling tiling;
    // device buffers
    double* xyz[3] = {nullptr, nullptr, nullptr};
    double* sintab[3] = {nullptr, nullptr, nullptr};  // sin-product MMS tables
    double* mms = nullptr;
    unsigned* fluid_mask = nullptr;
    int* fluid = nullptr;
    long nfluid = 0;
    int* faces = nullptr;
    long nfaces = 0;
    // Floating/Neumann bodies or Neumann grid faces: the reference schedule
    // only (cut-point rows, R sources, Neumann data); fused forms refuse
    bool general = false;
    bool any_free_r = false;      // some cut point of a non-Dirichlet body
    // the fused x-marching form also steps free cut points (no Neumann grid
    // faces): one record each, with the entries of its three derivatives'
    // rows in one list
    bool fused_free = false;
    shx::FreeRec* frees = nullptr;
    shx::FreeEnt* fents = nullptr;
    long nfree = 0;
    double* accF = nullptr;       // their fused accumulator register
    double* neuD = nullptr;       // heat: Neumann face data (N values)
    int* neu_faces = nullptr;     // points of the Neumann faces, face by face
    long neu_off[7] = {0, 0, 0, 0, 0, 0, 0};  // face f = 2 dir + side: [neu_off[f], neu_off[f+1])
    double* exR[3] = {nullptr, nullptr, nullptr};  // exact values at the cut points (statistics)
    double* rpos[3] = {nullptr, nullptr, nullptr};
    signed char* rdir[3] = {nullptr, nullptr, nullptr};
    // fused RK4 registers (stage ping-pong and accumulator), allocated lazily
    double* stageD[2] = {nullptr, nullptr};
    float* stageF[2] = {nullptr, nullptr};  // mixed form: FP32 stage register
    int precision = SHX_F64;
    double* stageR[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    double* accD = nullptr;
    // reference-schedule buffers, allocated lazily
    bool unfused_ready = false;
    double* gG[3][4] = {};
    double* du[3][4] = {};
    double* srcD = nullptr;
    shx_csr* Bfull[3] = {nullptr, nullptr, nullptr};
    // statistics partials
    double* part = nullptr;
    long* part_idx = nullptr;
    // fused-stage kernel choice: 0 point-wise tiles; 1-4 x-marching tiles
    // with the special work in a second launch (1), or in the same launch
    // before the tiles (2), after them (3) or spread among them (4)
    int variant = 0;
    int march_ty = 4;
    int march_xchunk = 16;
    int march_vec = 0;  // points per lane along z; 0: 2 when n2 is even, else 1
    // x-marching form: class byte per point of D, the special-point records
    // and the per-direction special-row tables
    unsigned char* cls = nullptr;
    shx::SpecialRec* special = nullptr;
    shx::SpecialRow* srows[3] = {nullptr, nullptr, nullptr};
    shx::ClosureRow* crows = nullptr;  // 3 x kMaxClosureRows
    int nrows[3] = {0, 0, 0};
    int fixw[3] = {0, 0, 0};  // every table row of a direction has `width` columns
    long nspecial = 0;
    long nface_items = 0;   // special items of the Dirichlet grid faces (face sizes, summed)
    long nface_points = 0;  // distinct points among them
    bool classes_ready = false;
    int width = 5;  // interior stencil width shared by the three derivatives
    shx_adjoint_ws* adj = nullptr;  // adjoint workspace, allocated on first use
    HostMms* hm = nullptr;          // SHX_MMS_HOST: the solution's functions and their sites
};

namespace shx {
void destroy_adjoint_ws(shx_adjoint_ws* w);
}  // namespace shx
