/*
 * operators.c -- oracle restatement of derivative assembly and application.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   csr::builder::to_csr         src/matrices/csr.cpp:10-35 (sort by (row, col, value),
 *                                duplicates kept, one row pointer per mesh point)
 *   csr::operator()              src/matrices/csr.cpp:37-51 (always +=, index order)
 *   submatrix_size               src/operators/derivative.cpp:259-303
 *   domain_discretization        src/operators/derivative.cpp:305-449
 *   inner_block placement        src/matrices/inner_block.cpp:17-40
 *   block device meta            src/matrices/block.hpp:30-94
 *   block matvec_functor         src/matrices/block.hpp:117-177 (per-row dot from 0.0,
 *                                left to right; rows positioned as the kernel does)
 *   derivative ctor / apply      src/operators/derivative.cpp:452-519
 *   gradient::operator()         src/operators/gradient.cpp:34-46
 *   laplacian::operator()        src/operators/laplacian.cpp:37-58
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

typedef struct {
    long row, col;
    double v;
} trip;

typedef struct {
    trip* p;
    long n, cap;
} tbuilder;

static void tb_add(tbuilder* b, long row, long col, double v)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->p = (trip*)realloc(b->p, (size_t)b->cap * sizeof(trip));
    }
    b->p[b->n].row = row;
    b->p[b->n].col = col;
    b->p[b->n].v = v;
    b->n++;
}

static int trip_cmp(const void* a, const void* b)
{
    const trip* x = (const trip*)a;
    const trip* y = (const trip*)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    if (x->v < y->v) return -1;
    if (x->v > y->v) return 1;
    return 0;
}

static void tb_to_csr(tbuilder* b, long nrows, orc_csr* A)
{
    qsort(b->p, (size_t)b->n, sizeof(trip), trip_cmp);
    A->rows = nrows;
    A->nnz = b->n;
    A->u = (long*)calloc((size_t)nrows + 1, sizeof(long));
    A->v = (long*)malloc((size_t)(b->n > 0 ? b->n : 1) * sizeof(long));
    A->w = (double*)malloc((size_t)(b->n > 0 ? b->n : 1) * sizeof(double));
    long first = 0;
    for (long i = 0; i < nrows; i++) {
        A->u[i + 1] = A->u[i];
        while (first < b->n && b->p[first].row == i) {
            ++A->u[i + 1];
            ++first;
        }
    }
    for (long k = 0; k < b->n; k++) {
        A->w[k] = b->p[k].v;
        A->v[k] = b->p[k].col;
    }
    free(b->p);
    b->p = NULL;
    b->n = b->cap = 0;
}

static long ic3(const orc_mesh* m, const int* ijk)
{
    return (long)ijk[0] * m->n[1] * m->n[2] + (long)ijk[1] * m->n[2] + ijk[2];
}

static void push_block(orc_deriv* d, long* cap, const orc_iblock* b)
{
    if (d->nblocks == *cap) {
        *cap = *cap ? 2 * *cap : 1024;
        d->blocks = (orc_iblock*)realloc(d->blocks, (size_t)*cap * sizeof(orc_iblock));
    }
    d->blocks[d->nblocks++] = *b;
}


/* ---------------------------------------------------------------- cut rows */

/* stencil::interp (stencil.hpp:213-270): the interpolant along dir at
 * closest + y between the line bounds; writes the bounds of the points it
 * spans to ol/or and returns its length */
static int stencil_interp(const orc_stencil* st, int dir, const int closest[3], double y,
                          const orc_bound* left, const orc_bound* right, double* c, int cap,
                          orc_bound* ol, orc_bound* orr)
{
    const orc_iinfo q = orc_query_interp(st);
    const int p = q.p, t = q.t;
    const int ic = closest[dir], lc = left->mesh[dir], rc = right->mesh[dir];
    int st1 = (p & 1) ? (p - 1) / 2 : p / 2;
    int st2 = (p - 1) / 2;
    if (y > 0) {
        const int tmp = st1;
        st1 = st2;
        st2 = tmp;
    }
    if ((ic - lc > st1 || (ic - lc == st1 && !left->has_obj)) &&
        (rc - ic > st2 || (rc - ic == st2 && !right->has_obj))) {
        memset(ol, 0, sizeof *ol);
        memset(orr, 0, sizeof *orr);
        for (int k = 0; k < 3; k++) ol->mesh[k] = orr->mesh[k] = closest[k];
        ol->mesh[dir] -= st1;
        orr->mesh[dir] += st2;
        return orc_interp_interior(st, y, c, cap);
    }
    if (ic - lc < rc - ic) {
        *ol = *left;
        memset(orr, 0, sizeof *orr);
        for (int k = 0; k < 3; k++) orr->mesh[k] = left->mesh[k];
        orr->mesh[dir] += t - 1;
        const double psi = left->has_obj ? left->psi : 1.0;
        if (left->has_obj && lc == ic) y += psi - 1;
        return orc_interp_wall(st, ic - lc, y, psi, c, 0, cap);
    }
    memset(ol, 0, sizeof *ol);
    for (int k = 0; k < 3; k++) ol->mesh[k] = right->mesh[k];
    ol->mesh[dir] -= t - 1;
    *orr = *right;
    const double psi = right->has_obj ? right->psi : 1.0;
    if (right->has_obj && rc == ic) y += 1 - psi;
    return orc_interp_wall(st, rc - ic, y, psi, c, 1, cap);
}

/* interp_deriv_coefficients (derivative.cpp:23-67): the derivative row along
 * dir used at a cut point of R(r), r != dir, and the offset of its first
 * point from the cut point's solid coordinate along dir */
static const double* interp_deriv_coeffs(int dir, int n, double h, const orc_rinfo* obj, int bc,
                                         const orc_stencil* st, double* c, double* ex, int* count,
                                         int* offset)
{
    const orc_info q = orc_query(st, bc);
    if (fabs(obj->normal[dir]) > 1e-12) {
        const int rw = obj->normal[dir] < 0.0;
        orc_nbs(st, h, bc, 1.0, rw, c, ex);
        *count = q.t;
        *offset = rw ? 1 - q.t : 0;
        return rw ? c + (q.r - 1) * q.t : c;
    }
    const int right_dist = n - obj->solid[dir], left_dist = obj->solid[dir];
    if (right_dist >= q.p && left_dist >= q.p) {
        orc_interior(st, h, c);
        *count = 2 * q.p + 1;
        *offset = -q.p;
        return c;
    }
    const int rw = right_dist < left_dist;
    orc_nbs(st, h, bc, 1.0, rw, c, ex);
    *count = q.t;
    if (rw) {
        const int qq = 1 + right_dist;
        *offset = qq - q.t;
        return c + (q.r - qq) * q.t;
    }
    *offset = -left_dist;
    return c + left_dist * q.t;
}

/* cut_discretization (derivative.cpp:159-257) with OB_builder (:69-157):
 * rows of R(r) for the derivative along dir, columns D (O) and R(r) (B) */
static void cut_discretization(int r, int dir, const orc_mesh* m, const orc_stencil* st,
                               const int* pbc, orc_csr* O, orc_csr* B)
{
    const long sz = m->nR[r];
    tbuilder Ob = {0}, Bb = {0};
    /* the reference's quick exit tests the shapes' conditions; skipping the
     * Dirichlet points one by one below gives the same rows either way */
    int all_dirichlet = 1;
    for (long k = 0; k < sz; k++) all_dirichlet &= pbc[k] == ORC_DIRICHLET;
    if (sz > 0 && !all_dirichlet) {
        const orc_info mx = orc_query_max(st);
        const double h = m->h[dir];
        double* c = (double*)calloc((size_t)(mx.r * mx.t + 16), sizeof(double));
        double* ic = (double*)calloc((size_t)(mx.t + 16), sizeof(double));
        double extra[16];
        const long stride = dir == 0 ? (long)m->n[1] * m->n[2] : dir == 1 ? (long)m->n[2] : 1;
        for (long row = 0; row < sz; row++) {
            const orc_rinfo* obj = &m->R[r][row];
            const int bc_t = pbc[row];
            if (bc_t == ORC_DIRICHLET) continue;
            if (dir == r) {
                /* the object's own closure row, from its solid coordinate */
                const orc_info q = orc_query(st, bc_t);
                orc_nbs(st, h, bc_t, obj->psi, obj->ray_outside, c, extra);
                const long sic = ic3(m, obj->solid);
                if (obj->ray_outside) {
                    const double* sub = c + (q.r - 1) * q.t; /* last row, reversed */
                    tb_add(&Bb, row, row, sub[q.t - 1]);
                    for (int i = 1; i < q.t; i++) tb_add(&Ob, row, sic - i * stride, sub[q.t - 1 - i]);
                } else {
                    tb_add(&Bb, row, row, c[0]);
                    for (int i = 1; i < q.t; i++) tb_add(&Ob, row, sic + i * stride, c[i]);
                }
                continue;
            }
            int count, off;
            const double* cl = interp_deriv_coeffs(dir, m->n[dir] - 1, h, obj, bc_t, st, c, extra, &count, &off);
            int cp[3] = {obj->solid[0], obj->solid[1], obj->solid[2]};
            if (obj->psi <= 0.5) cp[r] += 1 - 2 * obj->ray_outside;
            cp[dir] += off;
            const double y = (obj->ray_outside ? 1.0 : -1.0) * (obj->psi <= 0.5 ? obj->psi : obj->psi - 1);
            for (int k = 0; k < count; k++, ++cp[dir]) {
                const double v = cl[k];
                if (cp[dir] == obj->solid[dir]) {
                    tb_add(&Bb, row, row, v); /* add_cut_point */
                    continue;
                }
                orc_bound lb, rb, ol, orr;
                const long rs = orc_interp_line(m, r, cp, &lb, &rb);
                stencil_interp(st, r, cp, y, &lb, &rb, ic, mx.t, &ol, &orr);
                /* add_interp_row */
                const long lic = ic3(m, ol.mesh), ric = ic3(m, orr.mesh);
                int e = 0;
                if (ol.has_obj) tb_add(&Bb, row, ol.obj_coord, v * ic[e]);
                else tb_add(&Ob, row, lic, v * ic[e]);
                ++e;
                for (long x = lic + rs; x < ric; x += rs) tb_add(&Ob, row, x, v * ic[e++]);
                if (orr.has_obj) tb_add(&Bb, row, orr.obj_coord, v * ic[e]);
                else tb_add(&Ob, row, ric, v * ic[e]);
            }
        }
        free(c);
        free(ic);
    }
    tb_to_csr(&Ob, sz, O);
    tb_to_csr(&Bb, sz, B);
}

orc_deriv* orc_deriv_create(const orc_mesh* m, int dir, const orc_stencil* st,
                            const int grid_bc[6], const int* obj_bc)
{
    int* pbc[3];
    for (int r = 0; r < 3; r++) {
        pbc[r] = (int*)malloc((size_t)(m->nR[r] + 1) * sizeof(int));
        for (long k = 0; k < m->nR[r]; k++) pbc[r][k] = obj_bc[m->R[r][k].shape_id];
    }
    orc_deriv* d = orc_deriv_create_pbc(m, dir, st, grid_bc, (const int* const*)pbc);
    for (int r = 0; r < 3; r++) free(pbc[r]);
    return d;
}

orc_deriv* orc_deriv_create_pbc(const orc_mesh* m, int dir, const orc_stencil* st,
                                const int grid_bc[6], const int* const pbc[3])
{
    orc_deriv* d = (orc_deriv*)calloc(1, sizeof(orc_deriv));
    d->dir = dir;
    d->N = orc_mesh_size(m);
    if (m->n[dir] < 2) {
        d->B.u = (long*)calloc((size_t)d->N + 1, sizeof(long));
        d->B.rows = d->N;
        d->Nm.u = (long*)calloc((size_t)d->N + 1, sizeof(long));
        d->Nm.rows = d->N;
        for (int r = 0; r < 3; r++) {
            d->Bf[r].u = (long*)calloc((size_t)m->nR[r] + 1, sizeof(long));
            d->Bf[r].rows = m->nR[r];
            d->Br[r].u = (long*)calloc((size_t)m->nR[r] + 1, sizeof(long));
            d->Br[r].rows = m->nR[r];
        }
        return d;
    }

    orc_info mx = orc_query_max(st);
    const double h = m->h[dir];
    d->w = 2 * mx.p + 1;
    orc_interior(st, h, d->interior);

    double* left = (double*)calloc((size_t)(mx.r * mx.t + 8), sizeof(double));
    double* right = (double*)calloc((size_t)(mx.r * mx.t + 8), sizeof(double));
    double extra[8];
    tbuilder Bb = {0}, Nb = {0};
    long cap = 0;

    for (long li = 0; li < m->nlines[dir]; li++) {
        const orc_line* ln = &m->lines[dir][li];
        if (orc_dirichlet_line(m, ln->start.mesh, dir, grid_bc)) continue;
        const long stride = ln->stride;

        /* submatrix_size */
        long rows = ln->end.mesh[dir] - ln->start.mesh[dir] + 1;
        long columns = rows;
        long row_offset = ic3(m, ln->start.mesh);
        long col_offset = row_offset;
        const long last_row = ic3(m, ln->end.mesh);

        orc_iblock b;
        memset(&b, 0, sizeof b);

        if (ln->start.has_obj) {
            const int bc_t = pbc[dir][ln->start.obj_coord];
            orc_info q = orc_query(st, bc_t);
            orc_nbs(st, h, bc_t, ln->start.psi, 0, left, extra);
            int s = bc_t != ORC_DIRICHLET;
            int rL = q.r - s, tL = q.t;
            const double* lc = left + s * tL;
            b.left_rows = rL;
            b.left_cols = tL - 1;
            b.left = (double*)malloc((size_t)(rL * (tL - 1) + 1) * sizeof(double));
            for (int r = 0; r < rL; r++)
                for (int j = 0; j < tL - 1; j++) b.left[r * (tL - 1) + j] = lc[r * tL + 1 + j];
            /* remove_left_row_col */
            --rows;
            row_offset += stride;
            --columns;
            col_offset += stride;
            for (int r = 0; r < rL; r++)
                tb_add(&Bb, row_offset + stride * r, ln->start.obj_coord, lc[r * tL]);
        } else {
            const int bc = grid_bc[2 * dir];
            orc_info q = orc_query(st, bc);
            orc_nbs(st, h, bc, 1.0, 0, left, extra);
            b.left_rows = q.r;
            b.left_cols = q.t;
            b.left = (double*)malloc((size_t)(q.r * q.t + 1) * sizeof(double));
            memcpy(b.left, left, (size_t)(q.r * q.t) * sizeof(double));
            if (bc == ORC_DIRICHLET) {
                --rows;
                row_offset += stride;
                b.flags |= 1;
            } else if (bc == ORC_NEUMANN) {
                for (int r = 0; r < q.x; r++) tb_add(&Nb, row_offset + stride * r, row_offset, extra[r]);
            }
        }

        if (ln->end.has_obj) {
            const int bc_t = pbc[dir][ln->end.obj_coord];
            orc_info q = orc_query(st, bc_t);
            orc_nbs(st, h, bc_t, ln->end.psi, 1, right, extra);
            int s = bc_t != ORC_DIRICHLET;
            int rR = q.r - s, tR = q.t;
            const double* rc = right;
            b.right_rows = rR;
            b.right_cols = tR - 1;
            b.right = (double*)malloc((size_t)(rR * (tR - 1) + 1) * sizeof(double));
            for (int r = 0; r < rR; r++)
                for (int j = 0; j < tR - 1; j++) b.right[r * (tR - 1) + j] = rc[r * tR + j];
            /* remove_right_row_col */
            --rows;
            --columns;
            for (int r = 0; r < rR; r++)
                tb_add(&Bb, last_row + stride * (r - rR), ln->end.obj_coord, rc[r * tR + tR - 1]);
        } else {
            const int bc = grid_bc[2 * dir + 1];
            orc_info q = orc_query(st, bc);
            orc_nbs(st, h, bc, 1.0, 1, right, extra);
            b.right_rows = q.r;
            b.right_cols = q.t;
            b.right = (double*)malloc((size_t)(q.r * q.t + 1) * sizeof(double));
            memcpy(b.right, right, (size_t)(q.r * q.t) * sizeof(double));
            if (bc == ORC_DIRICHLET) {
                --rows;
                b.flags |= 2;
            } else if (bc == ORC_NEUMANN) {
                for (int r = 0; r < q.x; r++)
                    tb_add(&Nb, last_row + stride * (r - q.x + 1), last_row, extra[r]);
            }
        }

        b.interior_rows = rows - b.left_rows - b.right_rows;
        b.row_offset = row_offset;
        b.col_offset = col_offset;
        b.stride = stride;
        b.right_col_offset = col_offset + stride * (columns - b.right_cols);
        push_block(d, &cap, &b);
        /* A segment shorter than its two closure blocks would need a
         * circulant with a negative row count (derivative.cpp:428-436): its
         * closures overlap, which is undefined in the reference, so such
         * meshes are rejected; so are closures whose columns leave the mesh
         * buffer. */
        {
            long l1 = b.col_offset + (b.left_cols - 1) * stride;
            long r1 = b.right_col_offset + (b.right_cols - 1) * stride;
            if ((b.left_rows > 0 && (b.col_offset < 0 || l1 >= d->N)) ||
                (b.right_rows > 0 && (b.right_col_offset < 0 || r1 >= d->N)) || b.interior_rows < 0) {
                free(left);
                free(right);
                free(Bb.p);
                free(Nb.p);
                orc_deriv_destroy(d);
                return NULL;
            }
        }
    }

    tb_to_csr(&Bb, d->N, &d->B);
    tb_to_csr(&Nb, d->N, &d->Nm);
    free(left);
    free(right);
    /* derivative.cpp:470-472 */
    for (int r = 0; r < 3; r++) cut_discretization(r, dir, m, st, pbc[r], &d->Bf[r], &d->Br[r]);
    return d;
}

void orc_deriv_destroy(orc_deriv* d)
{
    if (!d) return;
    for (long i = 0; i < d->nblocks; i++) {
        free(d->blocks[i].left);
        free(d->blocks[i].right);
    }
    free(d->blocks);
    free(d->B.u);
    free(d->B.v);
    free(d->B.w);
    free(d->Nm.u);
    free(d->Nm.v);
    free(d->Nm.w);
    for (int r = 0; r < 3; r++) {
        free(d->Bf[r].u);
        free(d->Bf[r].v);
        free(d->Bf[r].w);
        free(d->Br[r].u);
        free(d->Br[r].v);
        free(d->Br[r].w);
    }
    free(d);
}

/* block.hpp:117-177 */
void orc_block_apply(const orc_deriv* d, const double* x, double* bout, int plus_eq)
{
    const long nb = d->nblocks;
    const int hw = d->w / 2;
#pragma omp parallel for schedule(static)
    for (long bi = 0; bi < nb; bi++) {
        const orc_iblock* m = &d->blocks[bi];
        const long total = m->left_rows + m->interior_rows + m->right_rows;
        for (long lr = 0; lr < total; lr++) {
            long out;
            double dot = 0;
            if (lr < m->left_rows) {
                long r = lr;
                out = m->row_offset + r * m->stride;
                for (long j = 0; j < m->left_cols; j++)
                    dot += m->left[r * m->left_cols + j] * x[m->col_offset + j * m->stride];
            } else if (lr < m->left_rows + m->interior_rows) {
                long r = lr - m->left_rows;
                out = m->row_offset + (m->left_rows + r) * m->stride;
                for (int j = 0; j < d->w; j++)
                    dot += d->interior[j] * x[out + (j - hw) * m->stride];
            } else {
                long r = lr - m->left_rows - m->interior_rows;
                out = m->row_offset + (m->left_rows + m->interior_rows + r) * m->stride;
                for (long j = 0; j < m->right_cols; j++)
                    dot += m->right[r * m->right_cols + j] * x[m->right_col_offset + j * m->stride];
            }
            if (plus_eq) bout[out] += dot;
            else bout[out] = dot;
        }
    }
}

/* csr.cpp:37-51 */
void orc_csr_apply(const orc_csr* A, const double* x, double* b)
{
    const long nr = A->rows;
#pragma omp parallel for schedule(static)
    for (long row = 0; row < nr; row++)
        for (long i = A->u[row]; i < A->u[row + 1]; i++) b[row] += A->w[i] * x[A->v[i]];
}

/* an R-space CSR apply: as csr::operator(), with column indices clamped to
 * [0, ncols) (a degenerate interpolation line of a scheme without
 * interpolants can name a point past the mesh; its weight is zero) */
static void csr_apply_clamped(const orc_csr* A, const double* x, long ncols, double* b)
{
    const long nr = A->rows;
#pragma omp parallel for schedule(static)
    for (long row = 0; row < nr; row++)
        for (long i = A->u[row]; i < A->u[row + 1]; i++) {
            long c = A->v[i];
            c = c < 0 ? 0 : (c >= ncols ? ncols - 1 : c);
            b[row] += A->w[i] * x[c];
        }
}

/* derivative.cpp:475-519: R rows (always +=: Bf from D, then Br from R), then
 * the D rows and their B terms */
void orc_deriv_apply(const orc_deriv* d, const orc_field* u, const orc_field* nu,
                     orc_field* du, int plus_eq)
{
    for (int r = 0; r < 3; r++)
        if (d->Bf[r].rows) csr_apply_clamped(&d->Bf[r], u->D, d->N, du->R[r]);
    for (int r = 0; r < 3; r++)
        if (d->Br[r].rows) csr_apply_clamped(&d->Br[r], u->R[r], d->Br[r].rows, du->R[r]);
    orc_block_apply(d, u->D, du->D, plus_eq);
    orc_csr_apply(&d->B, u->R[d->dir], du->D);
    if (nu) orc_csr_apply(&d->Nm, nu->D, du->D);
}
