/* euler.c -- CPU oracle of the compressible Euler system (test
 * infrastructure only; the product never calls it).
 *
 * The reference has no Euler system: src/systems/inviscid_vortex.cpp:118-154
 * is a stub whose rhs, update_boundary and initialize are empty.  This file
 * states the discretisation the product implements, built from the
 * reference's own pieces:
 *   - conserved variables U = (rho, rho u, rho v, rho w, E), ideal gas,
 *     p = (gamma - 1) (E - |m|^2 / (2 rho)) (gamma = 1.4, the stub's g);
 *   - rhs_k = -(D_x F_x,k + D_y F_y,k + D_z F_z,k), each D_d the first
 *     derivative operator of direction d (derivative::operator(), eq then
 *     plus_eq, derivative.cpp:475-519) applied to a flux field whose cut-point
 *     values enter through B (bodies are Dirichlet in the operators' sense);
 *   - slip wall at the bodies' cut points: the state of the fluid neighbour
 *     along the ray (solid point + 1, or - 1 when the ray leaves the body)
 *     with the normal momentum removed; density and pressure are kept;
 *   - Dirichlet grid faces hold their state (rhs zero there), floating faces
 *     take the operators' closures;
 *   - the RK4 slot schedule of rk4.cpp:16-47 over the five fields, the wall
 *     update after each stage input is formed.
 * Pinned by free-stream preservation, the isentropic vortex (the stub's dead
 * formulas, inviscid_vortex.cpp:13-116) and the wall-state invariants.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

struct orc_euler {
    const orc_mesh* m;
    const orc_deriv* d[3];
    double gamma;
    int grid_bc[6];
    long* wall_fluid[3]; /* per cut point of R(d): flat index of its fluid neighbour */
    orc_field F[5];      /* one direction's fluxes */
    orc_field du;
};

static void efield_alloc(const orc_mesh* m, orc_field* f)
{
    f->D = (double*)calloc((size_t)orc_mesh_size(m) + 1, sizeof(double));
    for (int d = 0; d < 3; d++) f->R[d] = (double*)calloc((size_t)m->nR[d] + 1, sizeof(double));
}

static void efield_free(orc_field* f)
{
    free(f->D);
    for (int d = 0; d < 3; d++) free(f->R[d]);
}

static void efield_zero(const orc_mesh* m, orc_field* f)
{
    memset(f->D, 0, (size_t)orc_mesh_size(m) * sizeof(double));
    for (int d = 0; d < 3; d++) memset(f->R[d], 0, (size_t)m->nR[d] * sizeof(double));
}

orc_euler* orc_euler_create(const orc_mesh* m, const orc_deriv* const d[3], double gamma,
                            const int grid_bc[6])
{
    orc_euler* e = (orc_euler*)calloc(1, sizeof(orc_euler));
    e->m = m;
    e->gamma = gamma;
    for (int c = 0; c < 3; c++) e->d[c] = d[c];
    for (int f = 0; f < 6; f++) e->grid_bc[f] = grid_bc[f];
    const long s[3] = {(long)m->n[1] * m->n[2], m->n[2], 1};
    for (int c = 0; c < 3; c++) {
        e->wall_fluid[c] = (long*)calloc((size_t)m->nR[c] + 1, sizeof(long));
        for (long i = 0; i < m->nR[c]; i++) {
            const orc_rinfo* r = &m->R[c][i];
            long idx = 0;
            for (int a = 0; a < 3; a++) idx += (long)r->solid[a] * s[a];
            e->wall_fluid[c][i] = idx + (r->ray_outside ? -s[c] : s[c]);
        }
    }
    for (int k = 0; k < 5; k++) efield_alloc(m, &e->F[k]);
    efield_alloc(m, &e->du);
    return e;
}

void orc_euler_destroy(orc_euler* e)
{
    if (!e) return;
    for (int c = 0; c < 3; c++) free(e->wall_fluid[c]);
    for (int k = 0; k < 5; k++) efield_free(&e->F[k]);
    efield_free(&e->du);
    free(e);
}

/* the five fluxes along direction dir at one state */
static void flux_at(double g, double r, double mx, double my, double mz, double E, int dir, double* F)
{
    const double p = (g - 1.0) * (E - 0.5 * (mx * mx + my * my + mz * mz) / r);
    const double m[3] = {mx, my, mz};
    const double ud = m[dir] / r;
    F[0] = m[dir];
    F[1] = mx * ud;
    F[2] = my * ud;
    F[3] = mz * ud;
    F[1 + dir] += p;
    F[4] = (E + p) * ud;
}

static void fluxes(const orc_euler* e, const orc_field* U, int dir, orc_field* F)
{
    const orc_mesh* m = e->m;
    const long N = orc_mesh_size(m);
    double f[5];
#pragma omp parallel for schedule(static) private(f)
    for (long i = 0; i < N; i++) {
        flux_at(e->gamma, U[0].D[i], U[1].D[i], U[2].D[i], U[3].D[i], U[4].D[i], dir, f);
        for (int k = 0; k < 5; k++) F[k].D[i] = f[k];
    }
    for (int c = 0; c < 3; c++)
        for (long i = 0; i < m->nR[c]; i++) {
            flux_at(e->gamma, U[0].R[c][i], U[1].R[c][i], U[2].R[c][i], U[3].R[c][i], U[4].R[c][i], dir, f);
            for (int k = 0; k < 5; k++) F[k].R[c][i] = f[k];
        }
}

void orc_euler_wall(const orc_euler* e, orc_field* U)
{
    const orc_mesh* m = e->m;
    const double g1 = e->gamma - 1.0;
    for (int c = 0; c < 3; c++)
        for (long i = 0; i < m->nR[c]; i++) {
            const long f = e->wall_fluid[c][i];
            const double* n = m->R[c][i].normal;
            const double r = U[0].D[f], E = U[4].D[f];
            double mm[3] = {U[1].D[f], U[2].D[f], U[3].D[f]};
            const double p = g1 * (E - 0.5 * (mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]) / r);
            const double mn = mm[0] * n[0] + mm[1] * n[1] + mm[2] * n[2];
            for (int a = 0; a < 3; a++) mm[a] = mm[a] - mn * n[a];
            U[0].R[c][i] = r;
            U[1].R[c][i] = mm[0];
            U[2].R[c][i] = mm[1];
            U[3].R[c][i] = mm[2];
            U[4].R[c][i] = p / g1 + 0.5 * (mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]) / r;
        }
}

void orc_euler_rhs(orc_euler* e, const orc_field* U, orc_field* rhs)
{
    const orc_mesh* m = e->m;
    const long N = orc_mesh_size(m);
    for (int k = 0; k < 5; k++) efield_zero(m, &rhs[k]);
    int first = 1;
    for (int dir = 0; dir < 3; dir++) {
        if (m->n[dir] < 2 || !e->d[dir]) continue;
        fluxes(e, U, dir, e->F);
        for (int k = 0; k < 5; k++) orc_deriv_apply(e->d[dir], &e->F[k], NULL, &rhs[k], !first);
        first = 0;
    }
    for (int k = 0; k < 5; k++) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; i++) rhs[k].D[i] = -rhs[k].D[i];
        for (int c = 0; c < 3; c++) memset(rhs[k].R[c], 0, (size_t)m->nR[c] * sizeof(double));
        /* Dirichlet grid faces hold their state */
        for (long i = 0; i < N; i++) {
            const long ijk[3] = {i / ((long)m->n[1] * m->n[2]), (i / m->n[2]) % m->n[1], i % m->n[2]};
            for (int a = 0; a < 3; a++)
                if ((e->grid_bc[2 * a] == ORC_DIRICHLET && ijk[a] == 0) ||
                    (e->grid_bc[2 * a + 1] == ORC_DIRICHLET && ijk[a] == m->n[a] - 1))
                    rhs[k].D[i] = 0.0;
        }
    }
}

static void lc(const orc_mesh* m, orc_field* d, const orc_field* s0, double c, const orc_field* r)
{
    const long N = orc_mesh_size(m);
    for (long i = 0; i < N; i++) d->D[i] = s0->D[i] + c * r->D[i];
    for (int a = 0; a < 3; a++)
        for (long i = 0; i < m->nR[a]; i++) d->R[a][i] = s0->R[a][i] + c * r->R[a][i];
}

static void acc(const orc_mesh* m, orc_field* d, double c, const orc_field* r)
{
    const long N = orc_mesh_size(m);
    for (long i = 0; i < N; i++) d->D[i] += c * r->D[i];
    for (int a = 0; a < 3; a++)
        for (long i = 0; i < m->nR[a]; i++) d->R[a][i] += c * r->R[a][i];
}

/* rk4.cpp:16-47 over the five fields */
void orc_euler_rk4_step(orc_euler* e, orc_field* u0, orc_field* out, orc_field* rk, orc_field* srhs,
                        double dt)
{
    static const double rki[4] = {0.0, 0.5, 0.5, 1.0};
    static const double rkf[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const orc_mesh* m = e->m;
    for (int k = 0; k < 5; k++) {
        efield_zero(m, &rk[k]);
        efield_zero(m, &srhs[k]);
        memcpy(out[k].D, u0[k].D, (size_t)orc_mesh_size(m) * sizeof(double));
        for (int a = 0; a < 3; a++) memcpy(out[k].R[a], u0[k].R[a], (size_t)m->nR[a] * sizeof(double));
    }
    for (int i = 0; i < 4; i++) {
        if (i > 0) {
            for (int k = 0; k < 5; k++) lc(m, &out[k], &u0[k], dt * rki[i], &srhs[k]);
            orc_euler_wall(e, out);
        }
        orc_euler_rhs(e, out, srhs);
        for (int k = 0; k < 5; k++) acc(m, &rk[k], dt * rkf[i], &srhs[k]);
    }
    for (int k = 0; k < 5; k++) lc(m, &out[k], &u0[k], 1.0, &rk[k]);
    orc_euler_wall(e, out);
}
