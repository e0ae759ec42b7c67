/*
 * systems.c -- oracle restatement of the scalar-wave and heat systems, the
 * gradient/laplacian drivers and the classical RK4 step.  TEST INFRASTRUCTURE
 * ONLY.
 *
 *   eval_at_locations           src/systems/detail/scalar_system_utils.hpp:16-73
 *   compute_scalar_stats        scalar_system_utils.hpp:77-181 (11-entry vector)
 *   initialize_scalar_field     scalar_system_utils.hpp:186-219
 *   scalar_wave ctor/rhs/bc     src/systems/scalar_wave.cpp:28-102,195-239,329-362
 *   heat rhs/fill_source/bc/dt  src/systems/heat.cpp:111-361
 *   gradient / laplacian        src/operators/gradient.cpp:34-46, laplacian.cpp:37-58
 *   rk4::operator()             src/temporal/rk4.cpp:11-47
 *   slot ops                    src/temporal/slot_ops.hpp:11-59
 *
 * The constant-velocity scalar wave (u_t + a.grad u = 0 with a translated
 * gaussian pulse as exact solution) is an extension the reference cannot
 * express; it uses the same gradient + dot path with gG = -a.  The heat
 * manufactured solution "sinprod" is the closed form sin(2 pi x) sin(2 pi y)
 * [sin(2 pi z)] exp(-t) that a Lua manufactured_solution table would supply.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double TWO_PI = 6.283185307179586; /* 2 * std::numbers::pi */

static double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

static void loc_of(const orc_mesh* m, long idx, double* loc)
{
    int ny = m->n[1], nz = m->n[2];
    long i = idx / ((long)ny * nz);
    long j = (idx / nz) % ny;
    long k = idx % nz;
    loc[0] = m->xyz[0][i];
    loc[1] = m->xyz[1][j];
    loc[2] = m->xyz[2][k];
}

/* ------------------------------------------------------- field helpers */
static void field_alloc(const orc_mesh* m, orc_field* f)
{
    f->D = (double*)calloc((size_t)orc_mesh_size(m), sizeof(double));
    for (int d = 0; d < 3; d++) f->R[d] = (double*)calloc((size_t)(m->nR[d] + 1), sizeof(double));
}

static void field_free(orc_field* f)
{
    free(f->D);
    for (int d = 0; d < 3; d++) free(f->R[d]);
}

static void field_zero(const orc_mesh* m, orc_field* f)
{
    memset(f->D, 0, (size_t)orc_mesh_size(m) * sizeof(double));
    for (int d = 0; d < 3; d++) memset(f->R[d], 0, (size_t)m->nR[d] * sizeof(double));
}

/* ------------------------------------------------------- exact solutions */
static double sw_exact(const orc_system* s, const double* loc, double t)
{
    if (s->kind == ORC_SYS_WAVE_RADIAL) {
        double r[3] = {loc[0] - s->center[0], loc[1] - s->center[1], loc[2] - s->center[2]};
        return sin(TWO_PI * (sqrt(dot3(r, r)) - s->radius - t));
    }
    double q[3];
    for (int c = 0; c < 3; c++) q[c] = loc[c] - s->x0[c] - s->vel[c] * t;
    return s->amp * exp(-dot3(q, q) / (2 * s->width * s->width));
}

static double sw_neg_g(const orc_system* s, int comp, const double* loc)
{
    if (s->kind == ORC_SYS_WAVE_RADIAL) {
        double r[3] = {loc[0] - s->center[0], loc[1] - s->center[1], loc[2] - s->center[2]};
        return -(loc[comp] - s->center[comp]) / sqrt(dot3(r, r));
    }
    return -s->vel[comp];
}

/* gauss3d (src/mms/gauss3d.cpp:10-111) restated term by term */
static double g3_expo(const double* g, const double* loc)
{
    const double *c = g, *v = g + 3;
    return (-0.5 * pow((loc[0] - c[0]), 2) * pow(v[0], -2) +
            (-0.5 * pow((loc[1] - c[1]), 2) * pow(v[1], -2) +
             -0.5 * pow((loc[2] - c[2]), 2) * pow(v[2], -2)));
}

/* the gaussian exponent of the mesh's dimension: gauss1d.cpp:14-16,
 * gauss2d.cpp:14-19, gauss3d (build_ms_gauss, mms.cpp:16-31) */
static double g_expo(const orc_system* s, const double* g, const double* loc)
{
    const double *c = g, *v = g + 3;
    if (s->mms_dims == 1) return -0.5 * pow((loc[0] - c[0]), 2) * pow(v[0], -2);
    if (s->mms_dims == 2)
        return (-0.5 * pow((loc[0] - c[0]), 2) * pow(v[0], -2) + -0.5 * pow((loc[1] - c[1]), 2) * pow(v[1], -2));
    return g3_expo(g, loc);
}

static double mms_call(const orc_system* s, double t, const double* loc)
{
    if (s->mms_kind == ORC_MMS_HOST) return s->fns.call(s->fns.ctx, t, loc);
    if (s->mms_kind == ORC_MMS_SINPROD) {
        double v = sin(TWO_PI * loc[0]) * sin(TWO_PI * loc[1]);
        if (s->mms_dims == 3) v = v * sin(TWO_PI * loc[2]);
        return v * exp(-t);
    }
    double sol = 0;
    for (int i = 0; i < s->mms_n; i++) {
        const double* g = s->mms + 8 * i;
        sol += exp(g_expo(s, g, loc)) * g[6] * cos(t * g[7]);
    }
    return sol;
}

static double mms_ddt(const orc_system* s, double t, const double* loc)
{
    if (s->mms_kind == ORC_MMS_HOST) return s->fns.ddt(s->fns.ctx, t, loc);
    if (s->mms_kind == ORC_MMS_SINPROD) return -mms_call(s, t, loc);
    double sol = 0;
    for (int i = 0; i < s->mms_n; i++) {
        const double* g = s->mms + 8 * i;
        sol += -exp(g_expo(s, g, loc)) * g[6] * g[7] * sin(t * g[7]);
    }
    return sol;
}

static void mms_grad(const orc_system* s, double t, const double* loc, double* out)
{
    out[0] = out[1] = out[2] = 0;
    if (s->mms_kind == ORC_MMS_HOST) {
        s->fns.grad(s->fns.ctx, t, loc, out);
        return;
    }
    if (s->mms_kind == ORC_MMS_SINPROD) {
        double sx = sin(TWO_PI * loc[0]), sy = sin(TWO_PI * loc[1]);
        double cx = cos(TWO_PI * loc[0]), cy = cos(TWO_PI * loc[1]);
        double sz = s->mms_dims == 3 ? sin(TWO_PI * loc[2]) : 1.0;
        double cz = s->mms_dims == 3 ? cos(TWO_PI * loc[2]) : 0.0;
        double e = exp(-t);
        out[0] = TWO_PI * cx * sy * sz * e;
        out[1] = TWO_PI * sx * cy * sz * e;
        out[2] = s->mms_dims == 3 ? TWO_PI * sx * sy * cz * e : 0.0;
        return;
    }
    for (int i = 0; i < s->mms_n; i++) {
        const double* g = s->mms + 8 * i;
        double base = -exp(g_expo(s, g, loc)) * g[6] * cos(t * g[7]);
        for (int d = 0; d < s->mms_dims; d++) out[d] += base * (loc[d] - g[d]) * pow(g[3 + d], -2);
    }
}

static double mms_lap(const orc_system* s, double t, const double* loc)
{
    if (s->mms_kind == ORC_MMS_HOST) return s->fns.lap(s->fns.ctx, t, loc);
    if (s->mms_kind == ORC_MMS_SINPROD)
        return -(s->mms_dims * TWO_PI * TWO_PI) * mms_call(s, t, loc);
    double sol = 0;
    const double x = loc[0], y = loc[1], z = loc[2];
    if (s->mms_dims == 1) { /* gauss1d.cpp:47-60 */
        for (int i = 0; i < s->mms_n; i++) {
            const double* c = s->mms + 8 * i;
            const double* v = c + 3;
            sol += c[6] * cos(t * c[7]) *
                   (exp(-0.5 * pow((x - c[0]), 2) * pow(v[0], -2)) * pow((x - c[0]), 2) * pow(v[0], -4) -
                    exp(-0.5 * pow((x - c[0]), 2) * pow(v[0], -2)) * pow(v[0], -2));
        }
        return sol;
    }
    if (s->mms_dims == 2) { /* gauss2d.cpp:65-85 */
        for (int i = 0; i < s->mms_n; i++) {
            const double* c = s->mms + 8 * i;
            const double* v = c + 3;
            sol += exp((-0.5 * pow((x - c[0]), 2) * pow(v[0], -2) + -0.5 * pow((y - c[1]), 2) * pow(v[1], -2))) *
                   c[6] * cos(t * c[7]) * pow(v[0], -4) * pow(v[1], -4) *
                   (((pow(y, 2) * pow(v[0], 4) +
                      (-2 * y * c[1] * pow(v[0], 4) +
                       (pow(c[1], 2) * pow(v[0], 4) +
                        (pow(x, 2) * pow(v[1], 4) +
                         (-2 * x * c[0] * pow(v[1], 4) + pow(c[0], 2) * pow(v[1], 4)))))) -
                     pow(v[0], 2) * pow(v[1], 4)) -
                    pow(v[0], 4) * pow(v[1], 2));
        }
        return sol;
    }
    for (int i = 0; i < s->mms_n; i++) {
        const double* c = s->mms + 8 * i;
        const double* v = c + 3;
        sol += exp(0.5 * ((-1 * pow((z - c[2]), 2) * pow(v[2], -2) -
                           pow((y - c[1]), 2) * pow(v[1], -2)) -
                          pow((x - c[0]), 2) * pow(v[0], -2))) *
               c[6] * cos(t * c[7]) * pow(v[0], -4) * pow(v[1], -4) * pow(v[2], -4) *
               ((((pow(z, 2) * pow(v[0], 4) * pow(v[1], 4) +
                   (-2 * z * c[2] * pow(v[0], 4) * pow(v[1], 4) +
                    (pow(c[2], 2) * pow(v[0], 4) * pow(v[1], 4) +
                     (pow(y, 2) * pow(v[0], 4) * pow(v[2], 4) +
                      (-2 * y * c[1] * pow(v[0], 4) * pow(v[2], 4) +
                       (pow(c[1], 2) * pow(v[0], 4) * pow(v[2], 4) +
                        (pow(x, 2) * pow(v[1], 4) * pow(v[2], 4) +
                         (-2 * x * c[0] * pow(v[1], 4) * pow(v[2], 4) +
                          pow(c[0], 2) * pow(v[1], 4) * pow(v[2], 4))))))))) -
                  pow(v[0], 2) * pow(v[1], 4) * pow(v[2], 4)) -
                 pow(v[0], 4) * pow(v[1], 2) * pow(v[2], 4)) -
                pow(v[0], 4) * pow(v[1], 4) * pow(v[2], 2));
    }
    return sol;
}

static double exact_at(const orc_system* s, const double* loc, double t)
{
    if (s->kind == ORC_SYS_HEAT) return mms_call(s, t, loc);
    return sw_exact(s, loc, t);
}

/* eval_at_locations(m, exact(t), out) */
void orc_system_exact(orc_system* s, orc_field* out, double t)
{
    const orc_mesh* m = s->m;
    const long N = orc_mesh_size(m);
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; idx++) {
        double loc[3];
        loc_of(m, idx, loc);
        out->D[idx] = exact_at(s, loc, t);
    }
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++) out->R[d][i] = exact_at(s, m->R[d][i].pos, t);
}

/* ---------------------------------------------- grid Dirichlet face lists */
typedef void (*face_fn)(long idx, void* ctx);

/* for_each_grid_bc_desc<Dirichlet> (selection_desc.hpp:211-222) */
static void for_each_dirichlet_face(const orc_system* s, face_fn fn, void* ctx)
{
    const orc_mesh* m = s->m;
    const int nx = m->n[0], ny = m->n[1], nz = m->n[2];
    for (int dir = 0; dir < 3; dir++) {
        for (int side = 0; side < 2; side++) {
            if (s->grid_bc[2 * dir + side] != ORC_DIRICHLET) continue;
            int p = side ? m->n[dir] - 1 : 0;
            if (dir == 0) {
                for (long q = 0; q < (long)ny * nz; q++) fn((long)p * ny * nz + q, ctx);
            } else if (dir == 1) {
                for (int i = 0; i < nx; i++)
                    for (int k = 0; k < nz; k++) fn((long)p * nz + (long)i * ny * nz + k, ctx);
            } else {
                for (long q = 0; q < (long)nx * ny; q++) fn(p + q * nz, ctx);
            }
        }
    }
}

static void zero_at(long idx, void* ctx) { ((double*)ctx)[idx] = 0.0; }

typedef struct {
    double* dst;
    const double* src;
} copy_ctx;

static void copy_at(long idx, void* ctx)
{
    copy_ctx* c = (copy_ctx*)ctx;
    c->dst[idx] = c->src[idx];
}

/* ------------------------------------------------------------ creation */
orc_system* orc_system_create(int kind, const int n[3], const double min[3],
                              const double max[3], int nshapes, const double* spheres,
                              const int* obj_bc, const int grid_bc[6],
                              const orc_stencil* st, const double* params, int nparams)
{
    double* sh = (double*)calloc((size_t)(nshapes > 0 ? nshapes : 1) * 8, sizeof(double));
    for (int q = 0; q < nshapes; q++) memcpy(sh + 8 * q + 1, spheres + 4 * q, 4 * sizeof(double));
    orc_system* s = orc_system_create8(kind, n, min, max, nshapes, sh, obj_bc, grid_bc, st, params, nparams);
    free(sh);
    return s;
}

orc_system* orc_system_create8(int kind, const int n[3], const double min[3],
                               const double max[3], int nshapes, const double* shapes8,
                               const int* obj_bc, const int grid_bc[6],
                               const orc_stencil* st, const double* params, int nparams)
{
    orc_system* s = (orc_system*)calloc(1, sizeof(orc_system));
    s->kind = kind;
    s->m = orc_mesh_create8(n, min, max, nshapes, shapes8);
    memcpy(s->grid_bc, grid_bc, sizeof s->grid_bc);
    s->obj_bc = (int*)malloc((size_t)(nshapes + 1) * sizeof(int));
    if (nshapes) memcpy(s->obj_bc, obj_bc, (size_t)nshapes * sizeof(int));
    const double* vel = kind == ORC_SYS_WAVE_LINEAR && nparams >= 8 ? params : NULL;
    for (int d = 0; d < 3; d++) {
        const orc_mesh* m = s->m;
        s->pbc[d] = (int*)malloc((size_t)(m->nR[d] + 1) * sizeof(int));
        for (long k = 0; k < m->nR[d]; k++) {
            const orc_rinfo* ri = &m->R[d][k];
            int bc = s->obj_bc[ri->shape_id];
            if (bc == ORC_INFLOW) {
                if (!vel) {
                    orc_system_destroy(s);
                    return NULL;
                }
                const double an = vel[0] * ri->normal[0] + vel[1] * ri->normal[1] + vel[2] * ri->normal[2];
                bc = an > 0.0 ? ORC_DIRICHLET : ORC_FLOATING;
            }
            s->pbc[d][k] = bc;
        }
    }
    for (int d = 0; d < 3; d++) {
        s->d[d] = orc_deriv_create_pbc(s->m, d, st, grid_bc, (const int* const*)s->pbc);
        if (!s->d[d]) {
            orc_system_destroy(s);
            return NULL;
        }
    }
    s->max_error = 100.0;
    if (kind == ORC_SYS_WAVE_RADIAL && nparams >= 4) {
        memcpy(s->center, params, 3 * sizeof(double));
        s->radius = params[3];
        if (nparams >= 5) s->max_error = params[4];
    } else if (kind == ORC_SYS_WAVE_LINEAR && nparams >= 8) {
        memcpy(s->vel, params, 3 * sizeof(double));
        memcpy(s->x0, params + 3, 3 * sizeof(double));
        s->width = params[6];
        s->amp = params[7];
        if (nparams >= 9) s->max_error = params[8];
    } else if (kind == ORC_SYS_HEAT && nparams >= 2) {
        s->diffusivity = params[0];
        s->mms_kind = (int)params[1];
        s->mms_dims = s->m->dims;
        if (s->mms_kind == ORC_MMS_GAUSS && nparams >= 3) {
            s->mms_n = (int)params[2];
            s->mms = (double*)malloc((size_t)(8 * s->mms_n + 1) * sizeof(double));
            memcpy(s->mms, params + 3, (size_t)(8 * s->mms_n) * sizeof(double));
        }
    } else {
        orc_system_destroy(s);
        return NULL;
    }
    for (int c = 0; c < 3; c++) {
        field_alloc(s->m, &s->gG[c]);
        field_alloc(s->m, &s->du[c]);
    }
    field_alloc(s->m, &s->src);
    field_alloc(s->m, &s->neu);

    if (kind != ORC_SYS_HEAT) {
        /* scalar_wave.cpp:70-95 */
        const orc_mesh* m = s->m;
        const long N = orc_mesh_size(m);
        for (int c = 0; c < 3; c++) {
            for (long idx = 0; idx < N; idx++) {
                double loc[3];
                loc_of(m, idx, loc);
                s->gG[c].D[idx] = sw_neg_g(s, c, loc);
            }
            for (int d = 0; d < 3; d++)
                for (long i = 0; i < m->nR[d]; i++) s->gG[c].R[d][i] = sw_neg_g(s, c, m->R[d][i].pos);
            for_each_dirichlet_face(s, zero_at, s->gG[c].D);
            for (int d = 0; d < 3; d++)
                for (long i = 0; i < m->nR[d]; i++)
                    if (s->pbc[d][i] == ORC_DIRICHLET) s->gG[c].R[d][i] = 0.0;
        }
    }
    return s;
}

/* the functions of an ORC_MMS_HOST manufactured solution (lua_mms.hpp:12-79:
 * call, ddt, grad, lap of (time, location)) */
void orc_system_set_mms(orc_system* s, const orc_mms_fns* f)
{
    s->fns = *f;
}

void orc_system_destroy(orc_system* s)
{
    if (!s) return;
    if (s->gG[0].D) {
        for (int c = 0; c < 3; c++) {
            field_free(&s->gG[c]);
            field_free(&s->du[c]);
        }
        field_free(&s->src);
        field_free(&s->neu);
    }
    for (int d = 0; d < 3; d++) orc_deriv_destroy(s->d[d]);
    orc_mesh_destroy(s->m);
    for (int d = 0; d < 3; d++) free(s->pbc[d]);
    free(s->obj_bc);
    free(s->mms);
    free(s);
}

orc_mesh* orc_system_mesh(orc_system* s) { return s->m; }
orc_deriv* orc_system_deriv(orc_system* s, int dir) { return s->d[dir]; }

/* ------------------------------------------------------ gradient / laplacian */
void orc_gradient_apply(const orc_system* s, const orc_field* u, orc_field* dux,
                        orc_field* duy, orc_field* duz)
{
    orc_field* du[3] = {dux, duy, duz};
    for (int c = 0; c < 3; c++) field_zero(s->m, du[c]);
    for (int c = 0; c < 3; c++)
        if (s->m->n[c] > 1) orc_deriv_apply(s->d[c], u, NULL, du[c], 0);
}

void orc_laplacian_apply(const orc_system* s, const orc_field* u, const orc_field* nu,
                         orc_field* du)
{
    field_zero(s->m, du);
    for (int c = 0; c < 3; c++)
        if (s->m->n[c] > 1) orc_deriv_apply(s->d[c], u, nu, du, 1);
}

/* ------------------------------------------------------------------- rhs */
void orc_system_rhs(orc_system* s, const orc_field* u, orc_field* rhs, double t)
{
    const orc_mesh* m = s->m;
    const long N = orc_mesh_size(m);
    if (s->kind != ORC_SYS_HEAT) {
        /* scalar_wave.cpp:195-239 */
        orc_gradient_apply(s, u, &s->du[0], &s->du[1], &s->du[2]);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; i++)
            rhs->D[i] = s->gG[0].D[i] * s->du[0].D[i] + s->gG[1].D[i] * s->du[1].D[i] +
                        s->gG[2].D[i] * s->du[2].D[i];
        for (int d = 0; d < 3; d++)
            for (long i = 0; i < m->nR[d]; i++)
                rhs->R[d][i] = s->gG[0].R[d][i] * s->du[0].R[d][i] +
                               s->gG[1].R[d][i] * s->du[1].R[d][i] +
                               s->gG[2].R[d][i] * s->du[2].R[d][i];
        return;
    }
    /* heat.cpp:119-162 */
    const double k = s->diffusivity;
    orc_laplacian_apply(s, u, &s->neu, rhs);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) rhs->D[i] *= k;
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++) rhs->R[d][i] *= k;
    /* fill_source (heat.cpp:111-117) */
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; idx++) {
        double loc[3];
        loc_of(m, idx, loc);
        s->src.D[idx] = mms_ddt(s, t, loc) - k * mms_lap(s, t, loc);
    }
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++) {
            const double* loc = m->R[d][i].pos;
            s->src.R[d][i] = mms_ddt(s, t, loc) - k * mms_lap(s, t, loc);
        }
    for (long q = 0; q < m->nfluid; q++) rhs->D[m->fluid[q]] += s->src.D[m->fluid[q]];
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++)
            if (s->pbc[d][i] != ORC_DIRICHLET) rhs->R[d][i] += s->src.R[d][i];
    for_each_dirichlet_face(s, zero_at, rhs->D);
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++)
            if (s->pbc[d][i] == ORC_DIRICHLET) rhs->R[d][i] = 0.0;
}

/* -------------------------------------------------------- update_boundary */
void orc_system_update_boundary(orc_system* s, orc_field* u, double t)
{
    const orc_mesh* m = s->m;
    orc_field sol;
    field_alloc(m, &sol);
    orc_system_exact(s, &sol, t);
    copy_ctx cc = {u->D, sol.D};
    for_each_dirichlet_face(s, copy_at, &cc);
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++)
            if (s->pbc[d][i] == ORC_DIRICHLET) u->R[d][i] = sol.R[d][i];
    if (s->kind == ORC_SYS_HEAT) {
        /* Neumann face data (heat.cpp:318-346) */
        for (int dir = 0; dir < 3; dir++) {
            for (int side = 0; side < 2; side++) {
                if (s->grid_bc[2 * dir + side] != ORC_NEUMANN) continue;
                int p = side ? m->n[dir] - 1 : 0;
                const long N = orc_mesh_size(m);
                for (long idx = 0; idx < N; idx++) {
                    long ijk[3] = {idx / ((long)m->n[1] * m->n[2]), (idx / m->n[2]) % m->n[1],
                                   idx % m->n[2]};
                    if (ijk[dir] != p) continue;
                    double loc[3], g[3];
                    loc_of(m, idx, loc);
                    mms_grad(s, t, loc, g);
                    s->neu.D[idx] = g[dir];
                }
            }
        }
    }
    field_free(&sol);
}

/* initialize_scalar_field (scalar_system_utils.hpp:186-219) */
void orc_system_initialize(orc_system* s, orc_field* u, double t)
{
    const orc_mesh* m = s->m;
    orc_field sol;
    field_alloc(m, &sol);
    orc_system_exact(s, &sol, t);
    memset(u->D, 0, (size_t)orc_mesh_size(m) * sizeof(double));
    for (long q = 0; q < m->nfluid; q++) u->D[m->fluid[q]] = sol.D[m->fluid[q]];
    for (int d = 0; d < 3; d++) memcpy(u->R[d], sol.R[d], (size_t)m->nR[d] * sizeof(double));
    field_free(&sol);
}

/* compute_scalar_stats (scalar_system_utils.hpp:77-181), serial first-max */
void orc_system_stats(orc_system* s, const orc_field* u, double t, double st[11])
{
    const orc_mesh* m = s->m;
    orc_field sol;
    field_alloc(m, &sol);
    orc_system_exact(s, &sol, t);
    /* MaxLoc starts from its identity (lowest value, largest int index), as
     * Kokkos 5.1 MaxLoc::init does; strict `>` keeps the first maximum and
     * never takes a NaN */
    double umin = DBL_MAX, umax = -DBL_MAX, err_d = -DBL_MAX;
    long err_idx = 2147483647L;
    for (long q = 0; q < m->nfluid; q++) {
        long i = m->fluid[q];
        if (u->D[i] < umin) umin = u->D[i];
        if (u->D[i] > umax) umax = u->D[i];
        double e = fabs(u->D[i] - sol.D[i]);
        if (e > err_d) {
            err_d = e;
            err_idx = i;
        }
    }
    if (m->nfluid == 0) umin = umax = err_d = 0.0;
    double ce[3] = {0, 0, 0}, ci[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++) {
        double best = -DBL_MAX;
        long bi = 2147483647L;
        int any = 0;
        for (long i = 0; i < m->nR[d]; i++) {
            if (s->pbc[d][i] == ORC_DIRICHLET) continue;
            any = 1;
            if (u->R[d][i] < umin) umin = u->R[d][i];
            if (u->R[d][i] > umax) umax = u->R[d][i];
            double e = fabs(u->R[d][i] - sol.R[d][i]);
            if (e > best) {
                best = e;
                bi = i;
            }
        }
        if (any) {
            ce[d] = best;
            ci[d] = (double)bi;
        }
    }
    double err = err_d;
    for (int d = 0; d < 3; d++)
        if (ce[d] > err) err = ce[d];
    st[0] = err;
    st[1] = umin;
    st[2] = umax;
    st[3] = err_d;
    st[4] = (double)err_idx;
    st[5] = ce[0];
    st[6] = ci[0];
    st[7] = ce[1];
    st[8] = ci[1];
    st[9] = ce[2];
    st[10] = ci[2];
    field_free(&sol);
}

/* heat.cpp:356-361, scalar_wave.cpp:357-362 */
double orc_system_dt(orc_system* s, double h_cfl, double p_cfl)
{
    double hmin = s->m->h[0];
    for (int d = 1; d < 3; d++)
        if (s->m->h[d] < hmin) hmin = s->m->h[d];
    if (s->kind == ORC_SYS_HEAT) return p_cfl * hmin * hmin / (4 * s->diffusivity);
    return h_cfl * hmin;
}


/* ----------------------------------------------------------------- rk4 */
static void slot_assign_lc(const orc_mesh* m, orc_field* dst, const orc_field* src, double c,
                           const orc_field* r)
{
    const long N = orc_mesh_size(m);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) dst->D[i] = src->D[i] + c * r->D[i];
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++) dst->R[d][i] = src->R[d][i] + c * r->R[d][i];
}

static void slot_accumulate(const orc_mesh* m, orc_field* dst, double c, const orc_field* r)
{
    const long N = orc_mesh_size(m);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) dst->D[i] += c * r->D[i];
    for (int d = 0; d < 3; d++)
        for (long i = 0; i < m->nR[d]; i++) dst->R[d][i] += c * r->R[d][i];
}

/* euler::operator() (src/temporal/euler.cpp:11-27): out <- u0, zero srhs,
 * srhs = RHS(out) at t, out = u0 + dt * srhs, boundary at t + dt */
void orc_euler_step(orc_system* s, orc_field* u0, orc_field* out, orc_field* srhs, double t,
                    double dt)
{
    const orc_mesh* m = s->m;
    memcpy(out->D, u0->D, (size_t)orc_mesh_size(m) * sizeof(double));
    for (int d = 0; d < 3; d++) memcpy(out->R[d], u0->R[d], (size_t)m->nR[d] * sizeof(double));
    field_zero(m, srhs);
    orc_system_rhs(s, out, srhs, t);
    slot_assign_lc(m, out, u0, dt, srhs);
    orc_system_update_boundary(s, out, t + dt);
}


void orc_rk4_step(orc_system* s, orc_field* u0, orc_field* out, orc_field* rk,
                  orc_field* srhs, double t, double dt)
{
    static const double rki[4] = {0.0, 0.5, 0.5, 1.0};
    static const double rkf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const orc_mesh* m = s->m;
    field_zero(m, rk);
    field_zero(m, srhs);
    memcpy(out->D, u0->D, (size_t)orc_mesh_size(m) * sizeof(double));
    for (int d = 0; d < 3; d++) memcpy(out->R[d], u0->R[d], (size_t)m->nR[d] * sizeof(double));
    for (int i = 0; i < 4; i++) {
        if (i > 0) {
            slot_assign_lc(m, out, u0, dt * rki[i], srhs);
            orc_system_update_boundary(s, out, t + dt * rki[i]);
        }
        orc_system_rhs(s, out, srhs, t + dt * rki[i]);
        slot_accumulate(m, rk, dt * rkf[i], srhs);
    }
    slot_assign_lc(m, out, u0, 1.0, rk);
    orc_system_update_boundary(s, out, t + dt);
}
