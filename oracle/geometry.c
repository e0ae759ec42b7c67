/*
 * geometry.c -- oracle restatement of the cartesian mesh, ray casting against
 * spheres, line segments and the fluid gather.  TEST INFRASTRUCTURE ONLY.
 *
 *   cartesian::cartesian        src/mesh/cartesian.cpp:13-44
 *   linear_distribute           src/fields/lazy_views.hpp:563-569
 *   sphere::hit / normal        src/mesh/sphere.cpp:15-45
 *   dot / length                src/real3_operators.hpp:62-84 (right fold)
 *   closest_hit, init_line<I>   src/mesh/object_geometry.cpp:17-116
 *   mesh init_line<I>           src/mesh/mesh.cpp:29-87
 *   init_slices, fluid gather   src/mesh/mesh.cpp:89-111,134-137;
 *                               src/fields/selection_desc.hpp:110-130
 *   mesh::dirichlet_line        src/mesh/mesh.cpp:154-168
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* index::dir<I>::slow / fast, src/indexing.hpp:21-37 */
static const int SLOW[3] = {1, 0, 0};
static const int FAST[3] = {2, 2, 1};

/* dot(u, v) is a unary right fold: u0*v0 + (u1*v1 + u2*v2) */
static double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

typedef struct {
    double t;
    double p[3];
    int ray_outside;
    int id;
} hit_t;

static int sphere_hit(const double* sph, int id, const double* o, const double* d,
                      double tmin, double tmax, hit_t* h)
{
    double oc[3] = {o[0] - sph[0], o[1] - sph[1], o[2] - sph[2]};
    double a = dot3(d, d);
    double b = dot3(oc, d);
    double c = dot3(oc, oc) - sph[3] * sph[3];
    double disc = b * b - a * c;
    if (disc > 0) {
        double sqr = sqrt(disc);
        for (int k = 0; k < 2; k++) {
            double t = k == 0 ? (-b - sqr) / a : (-b + sqr) / a;
            if (t > tmin && t < tmax) {
                double pc[3];
                h->t = t;
                for (int q = 0; q < 3; q++) h->p[q] = o[q] + t * d[q];
                for (int q = 0; q < 3; q++) pc[q] = h->p[q] - sph[q];
                h->ray_outside = dot3(d, pc) < 0;
                h->id = id;
                return 1;
            }
        }
    }
    return 0;
}

/* rect<I>::hit (src/mesh/rect.hpp:21-36): the plane normal to axis I at
 * rc[1], bounded by rc[2..3] - rc[4..5] in its (slow, fast) axes, fluid on
 * the side of rc[6].  Bounds are inclusive, as are t_min and t_max.  A line
 * lying in the plane (t = 0/0) takes no hit here; the reference computes
 * with a NaN t there. */
static int rect_hit(const double* rc, int id, const double* o, const double* d, double tmin,
                    double tmax, hit_t* h)
{
    const int I = (int)rc[0] - 1, S = SLOW[I], F = FAST[I];
    const double t = (rc[1] - o[I]) / d[I];
    if (t < tmin || t > tmax || isnan(t)) return 0;
    double p[3];
    for (int q = 0; q < 3; q++) p[q] = o[q] + t * d[q];
    if (p[S] < rc[2] || p[S] > rc[4] || p[F] < rc[3] || p[F] > rc[5]) return 0;
    h->t = t;
    memcpy(h->p, p, sizeof p);
    h->ray_outside = rc[6] * d[I] < 0;
    h->id = id;
    return 1;
}

/* object_geometry.cpp:17-30: the closest hit over the shapes in order */
static int closest_hit(const orc_mesh* m, const double* o, const double* d, double tmin,
                       double tmax, hit_t* out)
{
    int found = 0;
    for (int s = 0; s < m->nshapes; s++) {
        hit_t h;
        const double* sh = m->shapes8 + 8 * s;
        const int hit = sh[0] == 0 ? sphere_hit(sh + 1, s, o, d, tmin, tmax, &h)
                                   : rect_hit(sh, s, o, d, tmin, tmax, &h);
        if (hit) {
            *out = h;
            found = 1;
            tmax = h.t;
        }
    }
    return found;
}

static void shape_normal(const double* sh, const double* pos, double* n)
{
    if (sh[0] != 0) { /* rect<I>::normal (rect.hpp:38-43) */
        n[0] = n[1] = n[2] = 0.0;
        n[(int)sh[0] - 1] = sh[6];
        return;
    }
    const double* sph = sh + 1;
    double r[3] = {pos[0] - sph[0], pos[1] - sph[1], pos[2] - sph[2]};
    double len = sqrt(dot3(r, r));
    for (int q = 0; q < 3; q++) n[q] = r[q] / len;
}

static void push_r(orc_mesh* m, int I, long* cap, const orc_rinfo* r)
{
    if (m->nR[I] == *cap) {
        *cap = *cap ? 2 * *cap : 256;
        m->R[I] = (orc_rinfo*)realloc(m->R[I], (size_t)*cap * sizeof(orc_rinfo));
    }
    m->R[I][m->nR[I]++] = *r;
}

/* object_geometry.cpp:33-116 */
static void geometry_init_line(orc_mesh* m, int I)
{
    const int S = SLOW[I], F = FAST[I];
    long cap = 0;
    const double smin = m->min[S], sh = m->h[S];
    const double fmin = m->min[F], fh = m->h[F];
    const double min = m->min[I], max = m->max[I], h = m->h[I];
    for (int s = 0; s < m->n[S]; s++) {
        double origin[3] = {0, 0, 0};
        int coord[3] = {0, 0, 0};
        origin[S] = smin + s * sh;
        coord[S] = s;
        for (int f = 0; f < m->n[F]; f++) {
            origin[F] = fmin + f * fh;
            coord[F] = f;
            double tmin = 0;
            double tmax = max - min;
            origin[I] = min;
            double dirv[3] = {0, 0, 0};
            dirv[I] = 1.0;
            hit_t hit;
            while (closest_hit(m, origin, dirv, tmin, tmax, &hit)) {
                double t_over_h = hit.t / h;
                int i_cell = (int)round(t_over_h);
                const double snap_tol = 1e-12;
                if (fabs(t_over_h - i_cell) > snap_tol) i_cell = (int)t_over_h;
                coord[I] = i_cell + hit.ray_outside;
                int off = 1 - 2 * hit.ray_outside;
                double fluid_pos = min + h * (coord[I] + off);
                double psi = off * (fluid_pos - hit.p[I]) / h;
                if (psi < snap_tol) psi = snap_tol;
                if (psi > 1.0 - snap_tol) psi = 1.0 - snap_tol;
                orc_rinfo r;
                r.psi = psi;
                memcpy(r.pos, hit.p, sizeof r.pos);
                shape_normal(m->shapes8 + 8 * hit.id, hit.p, r.normal);
                r.ray_outside = hit.ray_outside;
                memcpy(r.solid, coord, sizeof r.solid);
                r.shape_id = hit.id;
                push_r(m, I, &cap, &r);
                tmin = nextafter(hit.t, tmax);
            }
        }
    }
}

static void push_line(orc_mesh* m, int I, long* cap, const orc_line* l)
{
    if (m->nlines[I] == *cap) {
        *cap = *cap ? 2 * *cap : 1024;
        m->lines[I] = (orc_line*)realloc(m->lines[I], (size_t)*cap * sizeof(orc_line));
    }
    m->lines[I][m->nlines[I]++] = *l;
}

static long ic3(const orc_mesh* m, const int* ijk)
{
    return (long)ijk[0] * m->n[1] * m->n[2] + (long)ijk[1] * m->n[2] + ijk[2];
}

/* mesh.cpp:29-87 */
static void mesh_init_line(orc_mesh* m, int I)
{
    if (m->n[I] == 1) return;
    const int S = SLOW[I], F = FAST[I];
    long stride = 1;
    for (int i = I + 1; i < 3; i++) stride *= m->n[i];
    long cap = 0;
    long first = 0, last = m->nR[I];
    const orc_rinfo* r = m->R[I];
    for (int s = 0; s < m->n[S]; s++) {
        for (int f = 0; f < m->n[F]; f++) {
            int left[3] = {0, 0, 0}, right[3] = {0, 0, 0};
            left[S] = right[S] = s;
            left[F] = right[F] = f;
            left[I] = 0;
            right[I] = m->n[I] - 1;
            int have_left = 1;
            orc_bound lb;
            memset(&lb, 0, sizeof lb);
            memcpy(lb.mesh, left, sizeof lb.mesh);
            while (first != last && left[S] == r[first].solid[S] &&
                   left[F] == r[first].solid[F]) {
                orc_bound ob;
                memcpy(ob.mesh, r[first].solid, sizeof ob.mesh);
                ob.has_obj = 1;
                ob.obj_coord = first;
                ob.obj_id = r[first].shape_id;
                ob.psi = r[first].psi;
                if (r[first].ray_outside) {
                    orc_line l;
                    l.stride = stride;
                    l.start = lb;
                    l.end = ob;
                    push_line(m, I, &cap, &l);
                    have_left = 0;
                } else {
                    lb = ob;
                    have_left = 1;
                }
                ++first;
            }
            if (have_left) {
                orc_line l;
                l.stride = stride;
                l.start = lb;
                memset(&l.end, 0, sizeof l.end);
                memcpy(l.end.mesh, right, sizeof l.end.mesh);
                push_line(m, I, &cap, &l);
            }
        }
    }
}

/* mesh.cpp:89-111 and selection_desc.hpp:110-130 */
static void mesh_init_fluid(orc_mesh* m)
{
    int i = m->n[2] > 1 ? 2 : m->n[1] > 1 ? 1 : 0;
    long ns = 0, cap = 1024;
    long* sl = (long*)malloc((size_t)cap * 2 * sizeof(long));
    for (long k = 0; k < m->nlines[i]; k++) {
        const orc_line* l = &m->lines[i][k];
        long i0 = l->start.has_obj ? ic3(m, l->start.mesh) + 1 : ic3(m, l->start.mesh);
        long i1 = l->end.has_obj ? ic3(m, l->end.mesh) : ic3(m, l->end.mesh) + 1;
        if (ns > 0 && sl[2 * (ns - 1) + 1] == i0) {
            sl[2 * (ns - 1) + 1] = i1;
        } else {
            if (ns == cap) {
                cap *= 2;
                sl = (long*)realloc(sl, (size_t)cap * 2 * sizeof(long));
            }
            sl[2 * ns] = i0;
            sl[2 * ns + 1] = i1;
            ns++;
        }
    }
    long total = 0;
    for (long k = 0; k < ns; k++) total += sl[2 * k + 1] - sl[2 * k];
    m->fluid = (long*)malloc((size_t)(total > 0 ? total : 1) * sizeof(long));
    long pos = 0;
    for (long k = 0; k < ns; k++)
        for (long idx = sl[2 * k]; idx < sl[2 * k + 1]; idx++) m->fluid[pos++] = idx;
    m->nfluid = total;
    free(sl);
}

orc_mesh* orc_mesh_create(const int n[3], const double min[3], const double max[3],
                          int nshapes, const double* spheres)
{
    double* sh = (double*)calloc((size_t)(nshapes > 0 ? nshapes : 1) * 8, sizeof(double));
    for (int s = 0; s < nshapes; s++) memcpy(sh + 8 * s + 1, spheres + 4 * s, 4 * sizeof(double));
    orc_mesh* m = orc_mesh_create8(n, min, max, nshapes, sh);
    free(sh);
    return m;
}

orc_mesh* orc_mesh_create8(const int n[3], const double min[3], const double max[3],
                           int nshapes, const double* shapes8)
{
    orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
    m->dims = 0;
    for (int i = 0; i < 3; i++) {
        m->n[i] = n[i] > 0 ? n[i] : 1;
        m->min[i] = min[i];
        m->max[i] = max[i];
        m->h[i] = (m->n[i] - 1) ? (m->max[i] - m->min[i]) / (m->n[i] - 1) : DBL_MAX;
        if (m->n[i] != 1) m->dims++;
        m->xyz[i] = (double*)malloc((size_t)m->n[i] * sizeof(double));
        for (int k = 0; k < m->n[i]; k++)
            m->xyz[i][k] = m->n[i] > 1
                               ? m->min[i] + k * (m->max[i] - m->min[i]) / (m->n[i] - 1)
                               : m->min[i];
    }
    if (nshapes < 0) nshapes = 0;
    m->nshapes = nshapes;
    m->shapes8 = (double*)malloc((size_t)(nshapes > 0 ? nshapes : 1) * 8 * sizeof(double));
    if (nshapes > 0) memcpy(m->shapes8, shapes8, (size_t)nshapes * 8 * sizeof(double));
    /* the spheres' {cx, cy, cz, r} (zero for other shapes) */
    m->spheres = (double*)calloc((size_t)(nshapes > 0 ? nshapes : 1) * 4, sizeof(double));
    for (int s = 0; s < nshapes; s++)
        if (shapes8[8 * s] == 0) memcpy(m->spheres + 4 * s, shapes8 + 8 * s + 1, 4 * sizeof(double));
    for (int I = 0; I < 3; I++) geometry_init_line(m, I);
    for (int I = 0; I < 3; I++) mesh_init_line(m, I);
    mesh_init_fluid(m);
    return m;
}

void orc_mesh_destroy(orc_mesh* m)
{
    if (!m) return;
    for (int i = 0; i < 3; i++) {
        free(m->xyz[i]);
        free(m->R[i]);
        free(m->lines[i]);
    }
    free(m->spheres);
    free(m->shapes8);
    free(m->fluid);
    free(m);
}

long orc_mesh_size(const orc_mesh* m) { return (long)m->n[0] * m->n[1] * m->n[2]; }

/* mesh.cpp:154-168 with cartesian::on_boundary (cartesian.hpp:88-92) */
int orc_dirichlet_line(const orc_mesh* m, const int start[3], int dir, const int bc[6])
{
    int result = 0;
    for (int i = 0; i < 3; i++) {
        if (i == dir) continue;
        int f = (bc[2 * i] == ORC_DIRICHLET && start[i] == 0) ||
                (bc[2 * i + 1] == ORC_DIRICHLET && start[i] == m->n[i] - 1);
        result = result || f;
    }
    return result;
}

/* mesh::interp_line (mesh.cpp:170-207): start from the grid ends of the line
 * through pt along dir and move them to the cut points of R(dir) on that line
 * that face pt (normal component >= 0 on the left, < 0 on the right). */
long orc_interp_line(const orc_mesh* m, int dir, const int pt[3], orc_bound* left, orc_bound* right)
{
    const int f = dir == 2 ? 1 : 2, s = dir == 0 ? 1 : 0; /* index::dirs (indexing.hpp:21-37) */
    memset(left, 0, sizeof *left);
    memset(right, 0, sizeof *right);
    for (int c = 0; c < 3; c++) left->mesh[c] = right->mesh[c] = pt[c];
    left->mesh[dir] = 0;
    right->mesh[dir] = m->n[dir] - 1;
    for (long i = 0; i < m->nR[dir]; i++) {
        const orc_rinfo* r = &m->R[dir][i];
        if (r->solid[s] > pt[s]) break;
        if (r->solid[s] != pt[s] || r->solid[f] != pt[f]) continue;
        orc_bound* b = NULL;
        if (r->normal[dir] >= 0.0 && r->solid[dir] <= pt[dir]) {
            if (r->solid[dir] >= left->mesh[dir]) b = left;
        } else if (r->normal[dir] < 0.0 && r->solid[dir] >= pt[dir]) {
            if (r->solid[dir] <= right->mesh[dir]) b = right;
        }
        if (b) {
            for (int c = 0; c < 3; c++) b->mesh[c] = r->solid[c];
            b->has_obj = 1;
            b->obj_coord = i;
            b->obj_id = r->shape_id;
            b->psi = r->psi;
        }
    }
    return dir == 0 ? (long)m->n[1] * m->n[2] : dir == 1 ? (long)m->n[2] : 1;
}
