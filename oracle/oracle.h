/*
 * oracle.h -- CPU restatement of the SHOCCS discrete-operator path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  It may be linked or loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  Nothing in
 * the product (shoccs_amd/) may include, link or call it.
 *
 * It restates, in plain C and in the reference's own evaluation order, the
 * functions listed below (file:line cites are relative to the reference tree):
 *
 *   geometry   src/mesh/cartesian.cpp:13-44, src/fields/lazy_views.hpp:563-569,
 *              src/mesh/sphere.cpp:15-45, src/mesh/object_geometry.cpp:17-116,
 *              src/mesh/mesh.cpp:29-111,154-168
 *   stencils   src/stencils/E4_1.cpp:66-342, E4u_1.cpp:49-126,
 *              E4_2.cpp:33-42,185-296, E2_2.cpp:95-175
 *   assembly   src/operators/derivative.cpp:259-449 (domain_discretization),
 *              src/matrices/csr.cpp:10-35 (builder), block.hpp:30-94 (meta)
 *   apply      src/matrices/block.hpp:117-177 (matvec_functor),
 *              src/matrices/csr.cpp:37-51, src/operators/derivative.cpp:475-519,
 *              src/operators/gradient.cpp:34-46, laplacian.cpp:37-58
 *   systems    src/systems/scalar_wave.cpp:28-362, heat.cpp:111-361,
 *              src/systems/detail/scalar_system_utils.hpp:16-219
 *   temporal   src/temporal/rk4.cpp:11-47, euler.cpp:11-27, slot_ops.hpp:11-59
 *
 * Arithmetic is compiled with -ffp-contract=off so that every sum and product
 * rounds exactly as the reference's x86-64 build does.
 *
 *   cut rows   src/operators/derivative.cpp:17-257 (interp_deriv_coefficients,
 *              OB_builder, cut_discretization), src/mesh/mesh.cpp:170-207
 *              (interp_line), src/stencils/stencil.hpp:213-270 (interp),
 *              E2_2.cpp:31-86, E4_2.cpp:31-183 (interpolants)
 */
#ifndef SHOCCS_ORACLE_H
#define SHOCCS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary condition types, src/operators/boundaries.hpp:13-20.  ORC_INFLOW
 * (no reference counterpart) is a shape condition of the constant-velocity
 * wave resolved per cut point: Dirichlet where the velocity leaves the body
 * into the fluid (a . n > 0, n the body's outward normal), Floating
 * elsewhere -- the reference's pattern of a Dirichlet upstream body and a
 * floating downstream body (scalar_wave.lua:10-22) applied point by point. */
enum { ORC_DIRICHLET = 0, ORC_FLOATING = 1, ORC_NEUMANN = 2, ORC_INFLOW = 3 };

/* stencil families, src/stencils/stencil.cpp:8-106 */
enum { ORC_E2_2 = 0, ORC_E4_2 = 1, ORC_E4_1 = 2, ORC_E4U_1 = 3, ORC_E2P_1 = 4 };

typedef struct { int p, r, t, x; } orc_info;
typedef struct { int p, t; } orc_iinfo; /* stencil.hpp:29 interp_info */
/* alpha: the scheme's free parameters (E4_1, E4u_1: 2; E2-poly: 13) */
typedef struct { int kind; double alpha[13]; } orc_stencil;

/* src/mesh/mesh_types.hpp:30-38 */
typedef struct {
    double psi;
    double pos[3];
    double normal[3];
    int ray_outside;
    int solid[3];
    int shape_id;
} orc_rinfo;

/* src/mesh/mesh_types.hpp:8-24 */
typedef struct {
    int mesh[3];
    int has_obj;
    long obj_coord;
    long obj_id;
    double psi;
} orc_bound;

typedef struct {
    long stride;
    orc_bound start, end;
} orc_line;

typedef struct {
    int n[3];
    double min[3], max[3], h[3];
    int dims;
    double* xyz[3];
    int nshapes;
    double* spheres; /* nshapes x {cx, cy, cz, r} */
    orc_rinfo* R[3];
    long nR[3];
    orc_line* lines[3];
    long nlines[3];
    long* fluid;
    long nfluid;
    /* shape records, 8 doubles each: sphere {0, cx, cy, cz, r}; rectangle
     * normal to axis a {1 + a, plane, c0_slow, c0_fast, c1_slow, c1_fast,
     * fluid_normal} (src/mesh/rect.hpp:11-44, rect.cpp:8-33) */
    double* shapes8;
} orc_mesh;

/* one inner_block in block-kernel (matvec_functor) terms, block.hpp:30-94 */
typedef struct {
    long row_offset, col_offset, stride;
    long left_rows, left_cols;
    double* left;
    long interior_rows;
    long right_rows, right_cols, right_col_offset;
    double* right;
    int flags; /* 1 = ldd, 2 = rdd (src/matrices/common.hpp:24-25) */
} orc_iblock;

typedef struct {
    long rows;
    long nnz;
    long* u; /* row pointer, rows + 1 */
    long* v; /* column */
    double* w;
} orc_csr;

typedef struct {
    int dir;
    long N;
    int w;
    double interior[9];
    orc_iblock* blocks;
    long nblocks;
    orc_csr B, Nm;
    /* R-space rows (cut_discretization, derivative.cpp:159-257): rows R(r),
     * columns D (Bf) and R(r) (Br); empty unless some body is not Dirichlet */
    orc_csr Bf[3], Br[3];
} orc_deriv;

/* a scalar field: D plus the three cut-point buffers, src/fields/scalar.hpp */
typedef struct {
    double* D;
    double* R[3];
} orc_field;

/* systems */
enum { ORC_SYS_WAVE_RADIAL = 0, ORC_SYS_WAVE_LINEAR = 1, ORC_SYS_HEAT = 2 };
/* manufactured solutions for heat */
enum { ORC_MMS_GAUSS = 0, ORC_MMS_SINPROD = 1, ORC_MMS_HOST = 2 };
/* a manufactured solution given by functions of (time, location), as the
 * reference's lua_mms (src/mms/lua_mms.hpp:12-79) calls its Lua functions */
typedef struct {
    void* ctx;
    double (*call)(void* ctx, double t, const double* loc);
    double (*ddt)(void* ctx, double t, const double* loc);
    void (*grad)(void* ctx, double t, const double* loc, double* g);
    double (*lap)(void* ctx, double t, const double* loc);
} orc_mms_fns;

typedef struct {
    int kind;
    orc_mesh* m;
    int grid_bc[6];
    int* obj_bc;
    int* pbc[3]; /* condition of each cut point of R(x), R(y), R(z) */
    orc_deriv* d[3];
    /* radial scalar wave (scalar_wave.cpp:28-40) */
    double center[3], radius;
    /* linear advection extension: constant velocity, gaussian pulse */
    double vel[3], x0[3], width, amp;
    double max_error;
    /* heat */
    double diffusivity;
    int mms_kind;
    int mms_dims;
    int mms_n;
    double* mms; /* per gaussian: center[3], variance[3], amplitude, frequency */
    orc_mms_fns fns; /* ORC_MMS_HOST */
    /* work buffers */
    orc_field gG[3];
    orc_field du[3];
    orc_field src;
    orc_field neu;
} orc_system;

/* compressible Euler (euler.c; no reference implementation, see there) */
typedef struct orc_euler orc_euler;

/* ------------------------------------------------------------------ API */
void orc_set_threads(int n);

orc_mesh* orc_mesh_create(const int n[3], const double min[3], const double max[3],
                          int nshapes, const double* spheres);
/* the same over shape records (orc_mesh.shapes8), spheres and rectangles */
orc_mesh* orc_mesh_create8(const int n[3], const double min[3], const double max[3],
                           int nshapes, const double* shapes8);
void orc_mesh_destroy(orc_mesh*);
long orc_mesh_size(const orc_mesh*);
int orc_dirichlet_line(const orc_mesh* m, const int start[3], int dir, const int bc[6]);

orc_info orc_query(const orc_stencil* st, int bc);
orc_info orc_query_max(const orc_stencil* st);
void orc_interior(const orc_stencil* st, double h, double* c);
void orc_nbs(const orc_stencil* st, double h, int bc, double psi, int right, double* c,
             double* extra);
/* interpolants: coefficients written to c, count returned (cap: the length of
 * c, returned unchanged by schemes without interpolants) */
orc_iinfo orc_query_interp(const orc_stencil* st);
int orc_interp_interior(const orc_stencil* st, double y, double* c, int cap);
int orc_interp_wall(const orc_stencil* st, int i, double y, double psi, double* c, int right, int cap);
/* mesh::interp_line (mesh.cpp:170-207): bounds of the line through pt along
 * dir; returns the stride */
long orc_interp_line(const orc_mesh* m, int dir, const int pt[3], orc_bound* left, orc_bound* right);

/* returns NULL on unsupported configuration; obj_bc: one condition per shape */
orc_deriv* orc_deriv_create(const orc_mesh* m, int dir, const orc_stencil* st,
                            const int grid_bc[6], const int* obj_bc);
/* the same with one condition per cut point (pbc[r][k] for point k of R(r)) */
orc_deriv* orc_deriv_create_pbc(const orc_mesh* m, int dir, const orc_stencil* st,
                                const int grid_bc[6], const int* const pbc[3]);
void orc_deriv_destroy(orc_deriv*);
void orc_block_apply(const orc_deriv* d, const double* x, double* b, int plus_eq);
void orc_csr_apply(const orc_csr* A, const double* x, double* b);
void orc_deriv_apply(const orc_deriv* d, const orc_field* u, const orc_field* nu,
                     orc_field* du, int plus_eq);

/* Non-Dirichlet bodies are accepted for the operators (derivative,
 * gradient, laplacian, with their cut-point rows); the system functions
 * (rhs, boundary update, stats, steps) restate the reference for Dirichlet
 * bodies only. */
orc_system* orc_system_create(int kind, const int n[3], const double min[3],
                              const double max[3], int nshapes, const double* spheres,
                              const int* obj_bc, const int grid_bc[6],
                              const orc_stencil* st, const double* params, int nparams);
/* the same over shape records (orc_mesh.shapes8) */
orc_system* orc_system_create8(int kind, const int n[3], const double min[3],
                               const double max[3], int nshapes, const double* shapes8,
                               const int* obj_bc, const int grid_bc[6],
                               const orc_stencil* st, const double* params, int nparams);
void orc_system_destroy(orc_system*);
void orc_system_set_mms(orc_system*, const orc_mms_fns*);
orc_mesh* orc_system_mesh(orc_system*);
orc_deriv* orc_system_deriv(orc_system*, int dir);
void orc_system_rhs(orc_system*, const orc_field* u, orc_field* rhs, double t);
void orc_system_update_boundary(orc_system*, orc_field* u, double t);
void orc_system_initialize(orc_system*, orc_field* u, double t);
void orc_system_exact(orc_system*, orc_field* out, double t);
void orc_system_stats(orc_system*, const orc_field* u, double t, double stats[11]);
double orc_system_dt(orc_system*, double h_cfl, double p_cfl);
void orc_euler_step(orc_system*, orc_field* u0, orc_field* out, orc_field* srhs, double t,
                    double dt);
void orc_rk4_step(orc_system*, orc_field* u0, orc_field* out, orc_field* rk,
                  orc_field* srhs, double t, double dt);

void orc_gradient_apply(const orc_system* s, const orc_field* u, orc_field* dux,
                        orc_field* duy, orc_field* duz);
void orc_laplacian_apply(const orc_system* s, const orc_field* u, const orc_field* nu,
                         orc_field* du);

#ifdef __cplusplus
}
#endif
/* Euler: fields are arrays of five orc_field (rho, rho u, rho v, rho w, E) */
orc_euler* orc_euler_create(const orc_mesh* m, const orc_deriv* const d[3], double gamma,
                            const int grid_bc[6]);
void orc_euler_destroy(orc_euler* e);
void orc_euler_wall(const orc_euler* e, orc_field* U);
void orc_euler_rhs(orc_euler* e, const orc_field* U, orc_field* rhs);
void orc_euler_rk4_step(orc_euler* e, orc_field* u0, orc_field* out, orc_field* rk, orc_field* srhs,
                        double dt);

#endif
