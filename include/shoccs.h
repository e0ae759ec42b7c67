/*
 * shoccs.h -- host-side C entry points of the SHOCCS MI355X library: build a
 * problem (mesh with embedded spheres, boundary conditions, stencil, system)
 * with the C++20 host library, inspect its assembled structures, and move
 * fields between host and device.  The compute calls themselves are the
 * shx_* functions of shoccs_hip.h, applied to the handles returned here.
 *
 * A problem replaces what the reference builds from a Lua `simulation` table
 * in system::from_lua (src/systems/system.cpp:103-132): mesh::from_lua
 * (src/mesh/mesh.cpp:209-220), bcs::from_lua (src/operators/boundaries.cpp:41-90),
 * stencil::from_lua (src/stencils/stencil.cpp:8-106) and the system ctor
 * (src/systems/heat.cpp:22-51, src/systems/scalar_wave.cpp:44-102).
 */
#ifndef SHOCCS_H
#define SHOCCS_H

#include <stdint.h>

#include "shoccs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* stencil schemes: scheme{order=2,type="E2"|"E4"}, scheme{order=1,type="E4"|"E4u"} */
#define SHOCCS_E2_2 0
#define SHOCCS_E4_2 1
#define SHOCCS_E4_1 2
#define SHOCCS_E4U_1 3
#define SHOCCS_E2P_1 4 /* "E2-poly" (src/stencils/polyE2_1.cpp) */

typedef struct shoccs_problem shoccs_problem;

typedef struct shoccs_problem_desc {
    int32_t kind;          /* SHX_SYS_WAVE_RADIAL, SHX_SYS_WAVE_LINEAR, SHX_SYS_HEAT */
    int32_t n[3];          /* index_extents (1 for unused axes) */
    double min[3], max[3]; /* domain_bounds */
    int32_t nshapes;
    const double* spheres; /* nshapes x {cx, cy, cz, r} (when shapes8 is NULL) */
    const int32_t* shape_bc; /* per shape boundary_condition */
    int32_t grid_bc[6];    /* xmin xmax ymin ymax zmin zmax */
    int32_t scheme;
    /* the scheme's free parameters: E4_1, E4u_1 two; E2-poly 13, its
     * floating (6), Dirichlet (3) and interpolant (4) parameters in order */
    double alpha[13];
    /* system parameters:
     *  WAVE_RADIAL: cx cy cz radius [max_error]
     *  WAVE_LINEAR: vx vy vz x0 y0 z0 width amplitude [max_error]
     *  HEAT:        diffusivity mms_kind [n_gauss, n_gauss x {c[3], var[3], amp, freq}]
     *               (mms_kind SHX_MMS_HOST: the functions in mms_fns) */
    const double* params;
    int32_t nparams;
    /* nonzero: the mesh and cut points only (no operators, no system), for
     * meshes whose lines are too short for the scheme's closures, e.g. the
     * field output of the reference's xdmf test (extents 3 4 5) */
    int32_t mesh_only;
    /* heat with mms_kind SHX_MMS_HOST: the manufactured solution's functions
     * (they must outlive the problem) */
    const shx_mms_fns* mms_fns;
    /* NULL, or nshapes shape records of 8 doubles in place of spheres:
     * a sphere {SHOCCS_SPHERE, cx, cy, cz, r, 0, 0, 0}; a rectangle normal to
     * axis a {SHOCCS_RECT_X + a, plane coordinate, c0_slow, c0_fast, c1_slow,
     * c1_fast, fluid_normal, 0} (src/mesh/rect.hpp:11-44, rect.cpp:8-33:
     * yz_rect, xz_rect, xy_rect; bounds in the plane's (slow, fast) axes,
     * y z for a yz rectangle; the fluid on the side fluid_normal points to) */
    const double* shapes8;
} shoccs_problem_desc;
#define SHOCCS_SPHERE 0
#define SHOCCS_RECT_X 1 /* yz_rect */
#define SHOCCS_RECT_Y 2 /* xz_rect */
#define SHOCCS_RECT_Z 3 /* xy_rect */

/* ctx == NULL builds the host structures only.  Problems with floating or
 * Neumann bodies, or Neumann grid faces, step with the reference schedule
 * (SHOCCS_RK4, SHOCCS_EULER); their fused form, batches and adjoint return
 * SHX_EUNSUPPORTED. */
int shoccs_problem_create(shx_ctx* ctx, const shoccs_problem_desc* desc, shoccs_problem** out);
/* The message of the calling thread's last failed host assembly
 * (shoccs_problem_create), or "" (the reference logs it and returns nullopt,
 * src/operators/derivative.cpp:452-473, system.cpp:103-131). */
const char* shoccs_last_error(void);
int shoccs_problem_destroy(shoccs_problem* p);
shx_system* shoccs_problem_system(shoccs_problem* p);
shx_derivative* shoccs_problem_derivative(shoccs_problem* p, int dir);
/* which = 0: mesh points N; 1, 2, 3: |Rx|, |Ry|, |Rz|; 4: fluid points */
int64_t shoccs_problem_size(const shoccs_problem* p, int which);
/* system::timestep_size (heat.cpp:356-361, scalar_wave.cpp:357-362) */
double shoccs_problem_dt(const shoccs_problem* p, double h_cfl, double p_cfl);

/* assembled structures, for bit-exact comparison with the reference's */
/* mesh extents and domain bounds (mesh::from_lua, src/mesh/cartesian.cpp:45-89) */
int shoccs_problem_geometry(const shoccs_problem* p, int32_t n[3], double lo[3], double hi[3]);
/* write_scalar_error (src/systems/detail/scalar_system_utils.hpp:224-301) on
 * host flat fields (D, Rx, Ry, Rz): err = |u - sol| at the fluid points and
 * at the cut points of non-Dirichlet bodies, 0 elsewhere and on Dirichlet
 * grid faces */
int shoccs_error_field(const shoccs_problem* p, const double* u, const double* sol, double* err);
/* outputs may be NULL */
int shoccs_mesh_r(const shoccs_problem* p, int dir, double* psi, double* pos, int32_t* solid,
                  int32_t* ray_outside, int32_t* shape_id);
int64_t shoccs_mesh_nlines(const shoccs_problem* p, int dir);
/* per line: start[dir], end[dir], start R index or -1, end R index or -1,
 * slow, fast, stride */
int shoccs_mesh_lines(const shoccs_problem* p, int dir, int64_t* out);
int shoccs_mesh_fluid(const shoccs_problem* p, int64_t* out);
int64_t shoccs_deriv_nblocks(const shoccs_problem* p, int dir);
int shoccs_deriv_blocks(const shoccs_problem* p, int dir, shx_line_block* out);
int64_t shoccs_deriv_npool(const shoccs_problem* p, int dir);
int shoccs_deriv_pool(const shoccs_problem* p, int dir, double* out);
int shoccs_deriv_interior(const shoccs_problem* p, int dir, double* out9);
/* The R-space rows of non-Dirichlet bodies (cut_discretization,
 * src/operators/derivative.cpp:159-257) in the layout of
 * shx_derivative_desc: ncut returns the row count and sets *nnz; cut copies
 * rset[rows], rrow[rows], rptr[rows + 1], rcol[nnz], rw[nnz]. */
int64_t shoccs_deriv_ncut(const shoccs_problem* p, int dir, int64_t* nnz);
int shoccs_deriv_cut(const shoccs_problem* p, int dir, int32_t* rset, int64_t* rrow, int64_t* rptr,
                     int64_t* rcol, double* rw);

/* ---------------------------------------------------------------- driver */
/* integrators of shoccs_run */
#define SHOCCS_RK4 0       /* rk4, the reference's slot schedule (shx_rk4_step) */
#define SHOCCS_RK4_FUSED 1 /* rk4, fused form at the system's precision and tuning */
#define SHOCCS_EULER 2     /* euler (shx_euler_step) */

typedef struct shoccs_run_desc {
    int32_t integrator;
    /* step_controller (src/temporal/step_controller.cpp:13-27): run while
     * step < max_step and time < max_time; both left at INT32_MAX / DBL_MAX
     * means zero steps.  dt = h_cfl h_min (wave) or p_cfl h_min^2 / 4k (heat);
     * a dt below min_dt ends the run prematurely. */
    int32_t max_step;
    double max_time;
    double h_cfl, p_cfl, min_dt;
    double* stats;     /* optional: up to stats_cap rows of the 11 statistics, */
    int64_t stats_cap; /* row 0 at t = 0, then one row per step */
    struct shoccs_io* io; /* optional field writer: U and Error at t = 0 and after
                           * each step, when its interval fires */
} shoccs_run_desc;

typedef struct shoccs_run_result {
    double time;       /* controller time at the end */
    int32_t step;      /* steps taken */
    int32_t completed; /* 1: the controller ran out; 0: ended prematurely (invalid
                        * statistics or dt < min_dt) */
    double linf;       /* the last L-inf error when completed, NaN otherwise */
    int64_t nstats;    /* rows written to stats */
} shoccs_run_result;

/* simulation_cycle::run (src/simulation/simulation_cycle.cpp:31-143):
 * initialize u at t = 0, update its boundary, then step while the
 * controller runs and system::valid(stats) holds (scalar_wave.cpp:108-112:
 * finite and |L-inf| <= max_error; heat.cpp:54-58: <= 1e6).  u holds the
 * final solution.  Synchronous. */
int shoccs_run(shoccs_problem* p, const shoccs_run_desc* d, shx_field* u, shoccs_run_result* r);

/* ------------------------------------------------------ Lua configuration
 * The reference's Lua `simulation` table (system::from_lua,
 * src/systems/system.cpp:103-132, and the from_lua readers it calls), read by
 * a Lua-subset evaluator (no Lua interpreter is linked): statements
 * `[local] name = exp`, `[local] function name (...) ... end`, call
 * statements; numbers, strings, booleans, tables, functions, arithmetic,
 * comparisons, `..`, and the math library; comments.  Function bodies run
 * locals, assignments, if/elseif/else and return.  Keys read:
 * mesh{index_extents, domain_bounds{min, max} | {...}},
 * domain_boundaries{xmin..zmax}, shapes{{type = "sphere", center, radius,
 * boundary_condition} | {type = "yz_rect", lower_corner, upper_corner, normal,
 * psi, boundary_condition}}, scheme{order, type, alpha} (E2/E4 order 2, E4/E4u
 * order 1), system{type = "heat" (diffusivity) | "scalar wave" (center,
 * radius, max_error) | "eigenvalues"}, manufactured_solution{type = "gaussian", {center,
 * variance, amplitude, frequency}...} | {type = "lua", call, ddt, grad,
 * div, lap} (functions of (time, loc), called on the host through
 * shx_mms_fns, src/mms/lua_mms.hpp:12-79), integrator{type = "rk4" | "euler"},
 * step_controller{max_step, max_time, min_dt, cfl{hyperbolic, parabolic}},
 * io{write_every_step, write_every_time, dir, suffix_length, xdmf_filename}
 * (shoccs_config_io); logging is ignored.  Extensions of this library:
 * system.velocity = {ax, ay, az} with system.initial_condition = {center,
 * width, amplitude} (constant-velocity wave with a gaussian pulse),
 * manufactured_solution.type = "sin_product", integrator.form = "fused".
 * On failure err (errlen bytes) holds the message. */
typedef struct shoccs_config shoccs_config;
int shoccs_config_parse(const char* text, shoccs_config** out, char* err, int64_t errlen);
/* the problem and run descriptors of a parsed configuration (the problem
 * descriptor's arrays belong to the configuration) */
int shoccs_config_desc(const shoccs_config* c, shoccs_problem_desc* desc, shoccs_run_desc* run);
/* 1 when the configuration's system.type is "eigenvalues"
 * (hyperbolic_eigenvalues.cpp: the problem then carries the mesh, conditions
 * and scheme on a constant-velocity wave along x that is never stepped; the
 * statistic -h_x min Re(eigenvalues of the x derivative) is
 * shoccs_amd.spectral.hyperbolic_eigenvalue_stat), else 0 */
int shoccs_config_eigenvalues(const shoccs_config* c);
/* the first error raised by a function of a "lua" manufactured solution
 * ("" when none; such an error makes the function's value NaN) */
const char* shoccs_config_error(const shoccs_config* c);
int shoccs_config_destroy(shoccs_config* c);

/* ------------------------------------------------------------ field output
 * field_io (src/io/field_io.cpp:32-65, field_data.cpp, xdmf.cpp): when the
 * dump interval fires (every write_every_step steps, every write_every_time
 * of simulated time within dt, and at step 0; interval.hpp:10-103), dump k
 * writes dir/<name>.<k, zero-padded to suffix_length> per variable (raw
 * native FP64: D, then Rx, Ry, Rz), the cut-point positions dir/rx, ry, rz at
 * the first dump (x, y, z per point; z, y, x on a mesh with one z point) and
 * the XDMF3 time series dir/<xdmf_filename>.  Neither interval set: nothing
 * is written. */
typedef struct shoccs_io shoccs_io;
typedef struct shoccs_io_desc {
    const char* dir;           /* io.dir; NULL or "": "io" */
    const char* xdmf_filename; /* io.xdmf_filename; NULL or "": "view.xmf" */
    int32_t suffix_length;     /* io.suffix_length; <= 0: 6 */
    int32_t write_every_step;  /* io.write_every_step; <= 0: unset */
    double write_every_time;   /* io.write_every_time; <= 0: unset */
} shoccs_io_desc;
int shoccs_io_create(const shoccs_problem* p, const shoccs_io_desc* d, shoccs_io** out);
int shoccs_io_destroy(shoccs_io* io);
/* host flat fields of the problem's shape; *wrote (may be NULL) = 1 when a
 * dump was written */
int shoccs_io_write(shoccs_io* io, int32_t nvars, const char* const* names, const double* const* fields,
                    int32_t step, double time, double dt, int32_t* wrote);
/* the configuration's io table (1 returned when it sets an interval, else
 * 0); the strings belong to the configuration */
int shoccs_config_io(const shoccs_config* c, shoccs_io_desc* io);

/* The compressible Euler system over a problem's mesh and first-derivative
 * operators (its bodies Dirichlet): the wall data (each cut point's fluid
 * neighbour along its ray and unit normal) from the mesh, then
 * shx_euler_create with the problem's grid boundary conditions. */
int shoccs_euler_create(shoccs_problem* p, double gamma, shx_euler** out);

/* device fields of the problem's shape: D then Rx, Ry, Rz */
int shoccs_field_alloc(shoccs_problem* p, shx_field* f);
int shoccs_field_free(shoccs_problem* p, shx_field* f);
int shoccs_field_upload(shoccs_problem* p, shx_field* f, const double* flat);
int shoccs_field_download(shoccs_problem* p, const shx_field* f, double* flat);

#ifdef __cplusplus
}
#endif
#endif
