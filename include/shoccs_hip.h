/*
 * shoccs_hip.h -- the device boundary of the SHOCCS discrete-operator path on
 * one MI355X (gfx950).
 *
 * Plain C ABI: device pointers, host pointers and sizes only.  Every function
 * returns 0 on success and a negative code on error; shx_last_error() returns
 * the message of the last failure on a context.  No exception crosses this
 * boundary.  All work of one context runs on that context's single HIP stream;
 * calls on one context must be serialized by the caller.  Entry points that
 * replace a synchronous reference call are synchronous unless their name ends
 * in _async.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference tree).  The reference is host-only C++ (Kokkos host
 * backend); it has no FFI of its own, so this header is the binding point a
 * maintainer would target from the reference's C++ classes.
 */
#ifndef SHOCCS_HIP_H
#define SHOCCS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHX_OK 0
#define SHX_EINVAL (-1)
#define SHX_EHIP (-2)
#define SHX_EUNSUPPORTED (-3)
#define SHX_ENOMEM (-4)

/* boundary condition types, src/operators/boundaries.hpp:13-20 */
#define SHX_BC_DIRICHLET 0
#define SHX_BC_FLOATING 1
#define SHX_BC_NEUMANN 2
/* A shape condition with no reference counterpart: each cut point of the
 * shape is Dirichlet where the constant velocity of the linear wave leaves
 * the body into the fluid (a . n > 0) and Floating elsewhere.  Only
 * shoccs_problem_create resolves it; operators and systems see the result. */
#define SHX_BC_INFLOW 3

typedef struct shx_ctx shx_ctx;
typedef struct shx_derivative shx_derivative;
typedef struct shx_csr shx_csr;
typedef struct shx_system shx_system;

/* A scalar field: the mesh buffer D (nx*ny*nz, z fastest) and the three
 * cut-point buffers Rx, Ry, Rz.  Replaces ccs::scalar_span / scalar_view
 * (src/fields/scalar.hpp:13-67).  All four are device pointers. */
typedef struct shx_field {
    double* D;
    double* R[3];
} shx_field;

/* One line segment of a derivative operator: a [dense | circulant | dense]
 * inner block placed along direction `dir` on the line identified by `pair`
 * (pair = slow * n_fast + fast, slow/fast as src/indexing.hpp:21-37).
 * Positions are coordinates along `dir`.  Replaces the per-line
 * inner_block_meta of src/matrices/inner_block_meta.hpp:9-23 with positions
 * instead of flat offsets, plus the B-matrix entries of the line's body
 * closures (src/operators/derivative.cpp:362-365,407-412) folded in. */
typedef struct shx_line_block {
    int32_t pair;  /* line identifier */
    int32_t lo;    /* position of the first row */
    int32_t nl;    /* left closure rows */
    int32_t ni;    /* circulant (interior) rows; negative when closures overlap */
    int32_t nr;    /* right closure rows */
    int32_t lcols; /* left closure columns */
    int32_t lcol0; /* position of the left closure's first column */
    int32_t lcoef; /* offset of the left closure (row-major) in the pool */
    int32_t rcols;
    int32_t rcol0;
    int32_t rcoef;
    int32_t lobj;  /* R(dir) index of the body at the left end, or -1 */
    int32_t lbw;   /* pool offset of the nl B weights of the left body */
    int32_t robj;  /* R(dir) index of the body at the right end, or -1 */
    int32_t rbw;   /* pool offset of the nr B weights of the right body */
    int32_t flags; /* bit0 ldd, bit1 rdd (src/matrices/common.hpp:24-25) */
} shx_line_block;

/* Everything the device needs for one direction's derivative. */
typedef struct shx_derivative_desc {
    int32_t dir;
    int32_t n[3];            /* mesh extents */
    int64_t nR[3];           /* |R(x)|, |R(y)|, |R(z)| */
    int32_t width;           /* interior stencil width 2p+1 (<= 9) */
    double interior[9];      /* interior stencil coefficients */
    const shx_line_block* blocks; /* sorted by (pair, lo) */
    int64_t nblocks;
    const double* pool;      /* closure coefficients and B weights */
    int64_t npool;
    /* Neumann rows, N matrix (derivative.cpp:375-379,422-426): row, col, w */
    const int64_t* neu_row;
    const int64_t* neu_col;
    const double* neu_w;
    int64_t nneu;
    /* R-space rows of non-Dirichlet bodies (derivative.cpp:159-257): row k is
     * row rrow[k] of R(rset[k]) with entries rptr[k]..rptr[k+1] (rptr has
     * nrrow + 1 values); an entry's column rcol >= 0 names a point of D,
     * rcol = -1 - c point c of R(rset[k]).  Within a row the D entries come
     * first; each row is summed in its stored order. */
    int64_t nrrow;
    const int32_t* rset;
    const int64_t* rrow;
    const int64_t* rptr;
    const int64_t* rcol;
    const double* rw;
} shx_derivative_desc;

/* ---------------------------------------------------------------- context */
int shx_init(int device, shx_ctx** ctx);
int shx_finalize(shx_ctx* ctx);
const char* shx_last_error(const shx_ctx* ctx);
int shx_sync(shx_ctx* ctx);
int shx_alloc(shx_ctx* ctx, size_t bytes, void** dptr);
int shx_free(shx_ctx* ctx, void* dptr);
int shx_h2d(shx_ctx* ctx, void* dst, const void* src, size_t bytes);
int shx_d2h(shx_ctx* ctx, void* dst, const void* src, size_t bytes);
/* dst[i] = value for i < n (a kernel, safe inside stream capture) */
int shx_fill(shx_ctx* ctx, double* dst, double value, int64_t n);
/* dst[i] = src[i] for i < n on the context's stream: one buffer of
 * field_registry::deep_copy_slot (src/fields/field_registry.hpp:172-187),
 * e.g. the u1 <- u0 copy of src/simulation/simulation_cycle.cpp:61. */
int shx_copy(shx_ctx* ctx, double* dst, const double* src, int64_t n);
/* Mean device time of shx_copy over `reps` launches, each bracketed by a HIP
 * event pair on the context's stream: the achievable-copy calibration the
 * roofline figures are read against. */
int shx_time_copy(shx_ctx* ctx, double* dst, const double* src, int64_t n, int reps,
                  double* ms_per_copy);

/* -------------------------------------------------------------- operators */
/* ccs::derivative::derivative (src/operators/derivative.cpp:452-473): upload
 * one direction's assembled operator. */
int shx_derivative_create(shx_ctx* ctx, const shx_derivative_desc* desc,
                          shx_derivative** out);
int shx_derivative_destroy(shx_derivative* d);

/* ccs::derivative::operator()(u, du, op) and (u, nu, du, op)
 * (src/operators/derivative.cpp:475-519): du.R[r] += the cut-point rows of
 * non-Dirichlet bodies over u.D and u.R[r] (always added, whatever the op),
 * du.D op= O u.D, du.D += B u.R[dir] (+= N nu.D when nu is given).
 * accumulate = 0 is eq_t, 1 is plus_eq_t.  Synchronous. */
int shx_derivative_apply(shx_derivative* d, const shx_field* u, const shx_field* nu,
                         shx_field* du, int accumulate);

/* ccs::matrix::block::operator()(x, b, op) (src/matrices/block.hpp:185-194):
 * the O part alone, every line of one direction in one launch. */
int shx_block_apply(shx_derivative* d, const double* x, double* b, int accumulate);

/* ccs::matrix::csr (src/matrices/csr.hpp:16-92, csr.cpp:10-51) in the
 * reference's own full-length layout (one row pointer per row); apply is
 * always b[row] += sum w*x[col]. */
int shx_csr_create(shx_ctx* ctx, int64_t rows, const int64_t* rowptr, const int64_t* col,
                   const double* w, shx_csr** out);
int shx_csr_destroy(shx_csr* A);
int shx_csr_apply(shx_csr* A, const double* x, double* b);

/* ccs::gradient::operator() (src/operators/gradient.cpp:34-46): zero the
 * three outputs (R parts included), then dx, dy, dz with eq_t on the
 * directions with extent > 1.  d[i] may be NULL for a direction of extent 1. */
int shx_gradient_apply(shx_derivative* const d[3], const shx_field* u, shx_field* const du[3]);

/* ccs::laplacian::operator() (src/operators/laplacian.cpp:37-58): zero du
 * (R parts included), then dx, dy, dz with plus_eq_t (and N nu when nu is
 * given); the cut-point rows of the three directions add up in du.R. */
int shx_laplacian_apply(shx_derivative* const d[3], const shx_field* u, const shx_field* nu,
                        shx_field* du);

/* ------------------------------------------------------ compressible Euler
 * No reference counterpart (src/systems/inviscid_vortex.cpp:118-154 is a
 * stub).  Five conserved fields (rho, rho u, rho v, rho w, E; arrays of five
 * shx_field below), ideal gas p = (gamma - 1)(E - |m|^2 / 2 rho);
 * rhs_k = -(D_x F_x,k + D_y F_y,k + D_z F_z,k) with the given first-derivative
 * operators (bodies Dirichlet in their sense: the cut points carry the wall
 * state); the slip wall sets each cut point to its fluid neighbour's state
 * with the normal momentum removed; Dirichlet grid faces hold their state. */
typedef struct shx_euler shx_euler;
typedef struct shx_euler_desc {
    double gamma;
    int32_t grid_bc[6];
    const int64_t* wall_fluid[3];  /* host, nR[d]: flat index of each cut point's fluid neighbour */
    const double* wall_normal[3];  /* host, 3 nR[d]: unit normal at each cut point */
} shx_euler_desc;
int shx_euler_create(shx_ctx* ctx, shx_derivative* const d[3], const shx_euler_desc* desc, shx_euler** out);
int shx_euler_destroy(shx_euler* e);
int shx_euler_wall(shx_euler* e, shx_field* U);
int shx_euler_rhs(shx_euler* e, const shx_field* U, shx_field* rhs);
/* rk4::operator() (src/temporal/rk4.cpp:16-47) over the five fields, the wall
 * update after each stage input and at the end; the result in out.  rk and
 * srhs are the schedule's scratch registers: srhs also holds stage inputs,
 * and the last stage's rk update is not stored.  Synchronous. */
int shx_euler_rk4_step(shx_euler* e, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs, double dt);
/* Precision of shx_euler_rk4_step (no reference counterpart).  SHX_F64 (the
 * default): plain FP64, identical to the CPU restatement.  SHX_F32_COMP: the
 * stage inputs after the first and the partial rhs are kept in FP32 at the
 * mesh points (two five-field FP32 buffers held by the system) and the
 * interior flux-derivative rows are computed and summed in FP32 with Kahan
 * compensation; closure rows, B terms, cut points, u0 and rk stay FP64, and
 * srhs is not used.  shx_euler_rhs stays FP64. */
int shx_euler_set_precision(shx_euler* e, int precision);
/* nsteps steps (u0 <- out after each), mean device time per step */
int shx_time_euler_steps(shx_euler* e, shx_field* u0, shx_field* out, shx_field* rk, shx_field* srhs, double dt,
                         int nsteps, double* ms_per_step);

/* ---------------------------------------------------------------- systems */
#define SHX_SYS_WAVE_RADIAL 0 /* src/systems/scalar_wave.cpp:28-40 */
#define SHX_SYS_WAVE_LINEAR 1 /* constant velocity, translated gaussian */
#define SHX_SYS_HEAT 2        /* src/systems/heat.cpp */
#define SHX_MMS_GAUSS 0       /* src/mms/gauss3d.cpp */
#define SHX_MMS_SINPROD 1     /* sin(2 pi x) sin(2 pi y) [sin(2 pi z)] exp(-t) */
#define SHX_MMS_HOST 2        /* functions of (time, location) called on the host (src/mms/lua_mms.hpp:12-79) */

/* A manufactured solution given by functions of time and location (x, y, z),
 * as the reference's lua_mms calls the Lua functions call, ddt, grad and lap
 * (src/mms/lua_mms.hpp:12-79, serial there too).  The heat system evaluates
 * them on the host, at the mesh and cut points that each use needs (initial
 * values, Dirichlet data, the source ddt - k lap, Neumann data, statistics),
 * and copies the values to the device; the functions must outlive the
 * system.  Such systems step with the reference schedule (shx_rk4_step,
 * shx_euler_step); fused forms, batches and the adjoint refuse them. */
typedef struct shx_mms_fns {
    void* ctx;
    double (*call)(void* ctx, double t, const double* loc);
    double (*ddt)(void* ctx, double t, const double* loc);
    void (*grad)(void* ctx, double t, const double* loc, double* g);
    double (*lap)(void* ctx, double t, const double* loc);
} shx_mms_fns;

typedef struct shx_system_desc {
    int32_t kind;
    int32_t n[3];
    const double* x; /* host coordinates, n[0], n[1], n[2] values */
    const double* y;
    const double* z;
    double h[3];
    int32_t grid_bc[6];   /* xmin xmax ymin ymax zmin zmax */
    int64_t nR[3];
    const double* r_pos[3];     /* host, 3 * nR[d] cut-point positions */
    const int32_t* r_dirichlet[3]; /* host, nR[d] flags: body is Dirichlet */
    const int64_t* fluid;       /* host fluid gather (mesh::fluid_desc) */
    int64_t nfluid;
    shx_derivative* d[3];       /* NULL for a direction of extent 1 */
    /* parameters */
    double center[3], radius;   /* radial wave */
    double vel[3], x0[3], width, amp; /* linear wave */
    double diffusivity;         /* heat */
    int32_t mms_kind, mms_dims, mms_n;
    const double* mms;          /* mms_n x {center[3], variance[3], amplitude, frequency} */
    const shx_mms_fns* mms_fns; /* SHX_MMS_HOST: the solution's functions */
} shx_system_desc;

int shx_system_create(shx_ctx* ctx, const shx_system_desc* desc, shx_system** out);
int shx_system_destroy(shx_system* s);

/* system::rhs / submit_rhs_graph (src/systems/system.cpp:34-69), fused:
 * scalar wave (scalar_wave.cpp:195-239): rhs = gG . grad u
 * heat (heat.cpp:119-162): rhs = k lap u + (S_t - k lap S) on the fluid, 0 on
 * Dirichlet faces and Dirichlet cut points.  One kernel for D. */
int shx_system_rhs(shx_system* s, const shx_field* u, shx_field* rhs, double t);

/* The same RHS in the reference's unfused schedule: zero fills, one block
 * kernel and one full-length CSR kernel per direction, then the dot /
 * scale / source / fill kernels.  Used as the plain FP64 GPU form. */
int shx_system_rhs_unfused(shx_system* s, const shx_field* u, shx_field* rhs, double t);

/* system::update_boundary (heat.cpp:296-354, scalar_wave.cpp:329-355):
 * Dirichlet grid faces of D and Dirichlet cut points take the exact solution. */
int shx_system_update_boundary(shx_system* s, shx_field* u, double t);
/* initialize_scalar_field (src/systems/detail/scalar_system_utils.hpp:186-219) */
int shx_system_initialize(shx_system* s, shx_field* u, double t);
/* eval_at_locations of the exact solution */
int shx_system_exact(shx_system* s, shx_field* out, double t);
/* compute_scalar_stats (scalar_system_utils.hpp:77-181): 11 doubles */
int shx_system_stats(shx_system* s, const shx_field* u, double t, double stats[11]);

/* integrators::rk4::operator() (src/temporal/rk4.cpp:16-47) with the
 * reference's slot schedule (zero rk/srhs, copy, 4 x {assign_lc,
 * update_boundary, rhs, accumulate}, final assign_lc + update_boundary), one
 * kernel per reference operation: the plain FP64 GPU form. */
int shx_rk4_step(shx_system* s, shx_field* u0, shx_field* out, shx_field* rk,
                 shx_field* srhs, double t, double dt);

/* integrators::euler (src/temporal/euler.cpp:11-27): out <- u0, zero srhs,
 * srhs = RHS(out) at t, out = u0 + dt srhs, update_boundary(out, t + dt). */
int shx_euler_step(shx_system* s, shx_field* u0, shx_field* out, shx_field* srhs, double t,
                   double dt);

/* The same step with the RHS, the stage update and the boundary update fused
 * into one kernel per stage over three registers (u0, stage, accumulator):
 * u0 is advanced in place by nsteps steps of size dt from time t.  Scratch
 * registers are owned by the system.  Asynchronous. */
int shx_rk4_fused_steps_async(shx_system* s, shx_field* u0, double t, double dt, int nsteps);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same steps as shx_rk4_fused_steps_async, bit for bit, with
 * the fused registers (accumulator, stage pair, stage cut-point values) taken
 * from a per-device workspace the library holds in its code object (64 MiB
 * per GPU, kept until the process ends) instead of the system's own buffers;
 * no hipMalloc, hipFree or other allocating call is made.  Limit: the
 * registers of the grid must fit the workspace (about 24 bytes per point in
 * FP64, 16 in the mixed form, plus 16 per cut point; about 2.7 million FP64
 * points); larger grids return SHX_EINVAL.  The x-marching form's point
 * classes must have been built beforehand (shx_system_stage_points, or any
 * fused step).  Every call on one GPU shares that GPU's workspace: calls that
 * overlap on one GPU (two streams, two contexts) race.  The registers need no
 * initial contents.  Asynchronous. */
int shx_rk4_fused_steps_noalloc(shx_system* s, shx_field* u0, double t, double dt, int nsteps);

/* Fused-stage kernel choice (no reference counterpart; the reference has one
 * Kokkos TeamPolicy shape, src/matrices/block.hpp:185-194): variant 0 is the
 * point-wise kernel on 64 x 4 tiles.  Variants 1-4 march tiles of
 * march_ty x 128 points (march_ty = 4 or 8; two points per lane, or
 * march_ty x 64 when n2 is odd) along x over chunks of
 * march_xchunk planes for interior, solid and wall-closure points, while
 * special workgroups evaluate the remaining points (closure rows with B
 * terms, closures out of the tile's reach, Dirichlet faces) from records
 * built once, and update the Dirichlet cut points: in a second launch (1),
 * or in the same launch before the tiles (2), after them (3) or spread among
 * them (4).  All variants give the same doubles.  A new system starts on
 * variant 3 (march_ty 4, march_xchunk 16) when its stencil width is 3 or 5
 * and N < 2^32. */
int shx_system_set_tuning(shx_system* s, int variant, int march_ty, int march_xchunk);
/* Precision of the fused form (no reference counterpart; the reference is
 * FP64 throughout).  SHX_F64: every register FP64, results bit-identical to
 * the reference's sums.  SHX_F32_COMP (the mixed form): the RK stage
 * register is stored in FP32 and interior stencils on it are summed in FP32
 * with Kahan compensation; u0, the RK accumulator, closure rows, B terms and
 * cut points stay FP64 (112 instead of 136 algorithmic bytes per point and
 * step).  Needs variants 1-4; selecting it on variant 0 switches to 3. */
#define SHX_F64 0
#define SHX_F32_COMP 1
int shx_system_set_precision(shx_system* s, int precision);
int shx_system_get_precision(const shx_system* s, int* precision);
int shx_system_get_tuning(const shx_system* s, int out[3]);
/* Points of D updated per stage by the timed fused-stage launch (out[0]) and,
 * among them, points the special-point work writes (out[1]: points with a
 * record and Dirichlet grid-face points; 0 for variant 0); builds the point
 * classes of variant 1 if needed. */
int shx_system_stage_points(shx_system* s, int64_t out[2]);
/* Autotuner: times every candidate on a copy of u0 (u0 is not modified),
 * keeps the fastest and returns its ms per step. */
int shx_system_autotune(shx_system* s, const shx_field* u0, double t, double dt, double* best_ms);

/* ------------------------------------------------------------------- batch */
/* Many independent problems stepped together (no reference counterpart; the
 * reference steps one problem per process): every fused RK4 stage of all of
 * them is one launch, their tile workgroups side by side, with per-problem
 * arguments and clocks (t = t0 + n dt) in device memory so that a captured
 * hipGraph stays valid at any time.  The systems must share the context,
 * system family, stencil width and precision, and take the x-marching form
 * (tile height march_ty for all; each keeps its x chunk).  u0[p] is advanced
 * in place; the batch keeps the field descriptors, not the memory. */
typedef struct shx_batch shx_batch;
int shx_batch_create(shx_ctx* ctx, shx_system* const* systems, const shx_field* u0,
                     const double* t0, const double* dt, int n, int march_ty, shx_batch** out);
int shx_batch_destroy(shx_batch* b);
/* nsteps steps of every problem (4 stage launches + 1 clock launch each) */
int shx_batch_steps_async(shx_batch* b, int nsteps);
/* capture nsteps steps as one hipGraph owned by the batch; each
 * shx_batch_launch_async replays it (advancing every problem nsteps steps) */
int shx_batch_capture(shx_batch* b, int nsteps);
int shx_batch_launch_async(shx_batch* b);
/* current time of every problem */
int shx_batch_times(shx_batch* b, double* t);

/* One problem's fused RK4 steps captured as a hipGraph (SURVEY's
 * shx_rk4_graph; the reference's rk4 has no graph form): nsteps steps from
 * t0 with step dt on u0 (kept by the graph; the system's fused registers and
 * tuning as shx_rk4_fused_steps_async uses them), the clock on the device.
 * Each shx_rk4_graph_launch replays the nsteps steps asynchronously on the
 * context's stream and advances the clock; the values equal nsteps calls of
 * shx_rk4_fused_steps_async.  Systems the batch refuses (Neumann grid faces,
 * host-evaluated solutions) are refused. */
typedef struct shx_rk4_graph shx_rk4_graph;
int shx_rk4_graph_create(shx_system* s, shx_field* u0, double t0, double dt, int nsteps, shx_rk4_graph** out);
int shx_rk4_graph_launch(shx_rk4_graph* g);
int shx_rk4_graph_time(shx_rk4_graph* g, double* t);
int shx_rk4_graph_destroy(shx_rk4_graph* g);

/* ----------------------------------------------------------------- adjoint
 * No reference counterpart: the reference has no adjoint.  Transposes of the
 * operators above and the discrete adjoint of the RK4 step, for gradients of a
 * final-time misfit with respect to the initial field and the interior
 * stencil coefficients. */

/* Transpose of one direction's derivative.  With du.D = O u.D + B u.R[dir]
 * the eq-form apply of shx_derivative_apply (no Neumann N term), sets
 * ubar.D = O^T dbar.D and ubar.R[dir] = B^T dbar.D (accumulate = 0), or adds
 * them (1).  Columns whose rows are all circulant rows take the reversed
 * stencil; every other column and every cut point sums its entries with rows
 * ascending.  Built on the first call (host, once per operator).  Operators
 * with cut-point rows (non-Dirichlet bodies) return SHX_EUNSUPPORTED. */
int shx_derivative_apply_T(shx_derivative* d, const shx_field* dbar, shx_field* ubar, int accumulate);

/* One fused FP64 RK4 step from u (read only) to out (a different field),
 * writing the stage inputs w1, w2, w3 (BU(u + dt rki[s] k_{s-1}) of
 * rk4.cpp:25-41, as the fused stages form them): the linearization points of
 * shx_rk4_adjoint_stages_async.  Asynchronous; same refusals as
 * shx_rk4_fused_steps_async, and FP64 only.  No reference counterpart. */
int shx_rk4_fused_keep_async(shx_system* s, const shx_field* u, shx_field* w1, shx_field* w2, shx_field* w3,
                             shx_field* out, double t, double dt);
/* The discrete adjoint of one RK4 step over its kept stage inputs
 * w[0..3] (w[0] the step's input, w[1..3] from shx_rk4_fused_keep_async): on
 * entry ubar holds dJ/du_{n+1}, on exit dJ/du_n, as shx_rk4_adjoint computes
 * it without recomputing the stages.  With with_coef the step's coefficient
 * gradient is added to a 27-entry accumulator on the device
 * (shx_adjoint_coef_set / shx_adjoint_coef_get, both synchronous).
 * Asynchronous.  No reference counterpart. */
int shx_rk4_adjoint_stages_async(shx_system* s, const shx_field* w, shx_field* ubar, double dt, int with_coef);
/* Builds the system's adjoint structures (workspace, the operators'
 * transposes, the merged closure lists), which the calls above otherwise
 * build on first use; synchronous. */
int shx_adjoint_prepare(shx_system* s);
int shx_adjoint_coef_set(shx_system* s, const double* g);
int shx_adjoint_coef_get(shx_system* s, double* g);

/* wbar = (d rhs / d u)^T kbar for the system's right-hand side (rhs rows on
 * D; the rows of the cut points are zero for Dirichlet bodies): scalar wave
 * sum_c D_c^T (gG_c kbar), heat sum_c D_c^T (k kbar), zero weight on
 * Dirichlet faces.  D and R parts of wbar are written. */
int shx_system_rhs_T(shx_system* s, const shx_field* kbar, shx_field* wbar);

/* Discrete adjoint of one RK4 step (rk4.cpp:16-47, FP64) taken from time t
 * with step dt: u holds the step's input u_n; on entry ubar holds
 * dJ/du_{n+1}, on exit dJ/du_n.  The stage inputs are recomputed from u_n;
 * entries that the boundary update overwrites (Dirichlet faces and cut
 * points) carry no sensitivity through it.  When coef_grad is not NULL, the
 * step's dJ/dcoef[c][m] (coef: the stored interior coefficients of direction
 * c's operator) is added to coef_grad[9 c + m] (host memory, 27 entries; the
 * sum runs on the device over the step's four stages, in stage order, and is
 * copied back once).  Synchronous. */
int shx_rk4_adjoint(shx_system* s, const shx_field* u, shx_field* ubar, double t, double dt,
                    double* coef_grad);

/* Interior (circulant) stencil coefficients of direction dir's operator:
 * replace them (width must match the operator's), or read them (up to 9). */
int shx_system_set_interior(shx_system* s, int dir, const double* coef, int width);
int shx_system_get_interior(const shx_system* s, int dir, double coef[9], int* width);

/* ------------------------------------------------------------- measurement */
/* Run `nsteps` fused steps `reps` times and report the mean device time per
 * step in milliseconds, measured with HIP events on the context's stream. */
int shx_time_fused_steps(shx_system* s, shx_field* u0, double t, double dt, int nsteps,
                         int reps, double* ms_per_step);
int shx_time_unfused_steps(shx_system* s, shx_field* u0, shx_field* out, shx_field* rk,
                           shx_field* srhs, double t, double dt, int nsteps,
                           double* ms_per_step);
/* Mean device time of the main fused stage kernel (see
 * shx_system_stage_points) over `nsteps` fused steps, each launch bracketed
 * by a HIP event pair on the context's stream. */
int shx_time_fused_stage_kernel(shx_system* s, shx_field* u0, double t, double dt, int nsteps,
                                double* ms_per_kernel);
/* Mean device time of one derivative apply (O and B, no Neumann term) over
 * `reps` applies after one warm-up, as the reference's derivative benchmark
 * times derivative::operator() (benchmarks/bench_derivative.cpp:56-116). */
int shx_time_derivative_apply(shx_derivative* d, const shx_field* u, shx_field* du, int accumulate,
                              int reps, double* ms_per_apply);
/* Mean device time per step of `reps` batch steps (use_graph = 0) or graph
 * replays (use_graph = 1, per captured step); points = total points of D. */
int shx_time_batch(shx_batch* b, int use_graph, int reps, double* ms_per_step, int64_t* points);

#ifdef __cplusplus
}
#endif
#endif
