// shx_state.hpp -- host-side state behind the opaque handles of shoccs_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "shoccs_hip.h"
#include "shx_internal.hpp"

struct shx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // this device's address of the allocation-free workspace (shx_core.hip),
    // or nullptr when it could not be resolved
    void* scratch = nullptr;
};

struct shx_deriv_T;  // transposed operator (shx_adjoint.hip), built on first use

struct shx_derivative {
    shx_ctx* ctx = nullptr;
    int dir = 0;
    int n[3] = {1, 1, 1};
    long nR[3] = {0, 0, 0};
    shx::DevDeriv dev;
    int* pair_ptr = nullptr;
    unsigned long long* fast = nullptr;
    long nslow = 0;                  // rows outside the fast intervals
    long nslow_list = 0;             // those not in a WallSide: the first of `slow`
    shx::SlowRow* slow = nullptr;    // sorted by point (the first nslow_list, then the rest)
    long nwall = 0;                  // closure ends evaluated one lane each
    shx::WallSide* walls = nullptr;  // sorted by first row position, then line
    shx_line_block* blocks = nullptr;
    double* pool = nullptr;
    double* coef = nullptr;
    long nneu_rows = 0;
    long* neu_rows = nullptr;
    int* neu_ptr = nullptr;
    long* neu_col = nullptr;
    double* neu_w = nullptr;
    // R-space rows of non-Dirichlet bodies (shx_derivative_desc.rrow ...)
    long nrrow = 0;
    int* rset = nullptr;
    long* rrow = nullptr;
    long* rptr = nullptr;
    long* rcol = nullptr;
    double* rw = nullptr;
    // host copies kept to build the reference's full-length B matrix and the
    // special-point rows of the x-marching form
    std::vector<shx_line_block> h_blocks;
    std::vector<double> h_pool;
    std::vector<int> h_pair_ptr;
    std::vector<int> h_rset;   // R-space rows: set and row of each, and
    std::vector<long> h_rrow;  // their entries
    std::vector<long> h_rptr, h_rcol;
    std::vector<double> h_rw;
    double h_coef[9] = {};  // interior stencil
    shx_deriv_T* T = nullptr;
};

struct shx_csr {
    shx_ctx* ctx = nullptr;
    long rows = 0;
    long nnz = 0;
    long* u = nullptr;
    long* v = nullptr;
    double* w = nullptr;
};

namespace shx {

// Size of the per-device workspace behind the allocation-free entry points
// (shx_rk4_fused_steps_noalloc): 64 MiB, i.e. the fused FP64 registers of a
// grid of up to about 2.7 million points.
constexpr size_t kScratchBytes = size_t(64) << 20;

int fail(shx_ctx* ctx, int code, const std::string& msg);
int hip_fail(shx_ctx* ctx, hipError_t e, const char* what);
int launch_fill(shx_ctx* ctx, double* dst, double v, long n);
int launch_derivative(shx_derivative* d, const double* u, const double* uR, double* du, bool plus,
                      bool with_b);
int launch_neumann(shx_derivative* d, const double* nu, double* du);
int launch_cut_rows(shx_derivative* d, const shx_field* u, const shx_field* du);
int launch_csr_full(shx_csr* A, const double* x, double* b);
int zero_field(shx_ctx* ctx, const shx_field* f, const int n[3], const long nR[3]);
int sync(shx_ctx* ctx);
void destroy_transpose(shx_deriv_T* T);
template <typename T>
int upload(shx_ctx* ctx, T** dst, const T* src, size_t n);

}  // namespace shx
