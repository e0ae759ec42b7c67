// problem.cpp -- C entry points of the host library (include/shoccs.h): build
// a mesh, its boundary conditions, stencil and system, upload them through
// shoccs_hip.h, and expose the assembled structures.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "operators.hpp"
#include "shoccs.h"

namespace ccs {

struct problem {
    shx_ctx* ctx = nullptr;
    int kind = 0;
    std::unique_ptr<mesh> m;
    bcs::Grid grid;
    bcs::Object obj;
    bcs::Points pbc;  // the condition of every cut point
    std::unique_ptr<derivative> d[3];
    shx_system* sys = nullptr;
    double diffusivity = 1.0;
    double max_error = 100.0;  // system::valid bound
};

}  // namespace ccs

struct shoccs_problem : ccs::problem {
};

using namespace ccs;

namespace {
thread_local std::string g_last_error;

int host_fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}
}  // namespace

extern "C" {

const char* shoccs_last_error(void) { return g_last_error.c_str(); }

int shoccs_problem_create(shx_ctx* ctx, const shoccs_problem_desc* desc, shoccs_problem** out)
{
    if (!desc || !out) return SHX_EINVAL;
    *out = nullptr;
    g_last_error.clear();
    const double* pr = desc->params;
    const int np = desc->nparams;
    if (desc->kind == SHX_SYS_WAVE_RADIAL && np < 4) return SHX_EINVAL;
    if (desc->kind == SHX_SYS_WAVE_LINEAR && np < 8) return SHX_EINVAL;
    if (desc->kind == SHX_SYS_HEAT && np < 2) return SHX_EINVAL;
    auto st = stencil::make(desc->scheme, desc->alpha, 13);
    if (!st) return SHX_EINVAL;

    auto p = std::unique_ptr<shoccs_problem>(new (std::nothrow) shoccs_problem());
    if (!p) return SHX_ENOMEM;
    p->ctx = ctx;
    p->kind = desc->kind;
    std::vector<body_shape> shapes;
    for (int s = 0; s < desc->nshapes; ++s) {
        body_shape b;
        if (desc->shapes8) {
            const double* q = desc->shapes8 + 8 * s;
            const int kind = (int)q[0];
            if (kind < SHOCCS_SPHERE || kind > SHOCCS_RECT_Z || (double)kind != q[0])
                return host_fail(SHX_EINVAL, "shape " + std::to_string(s) + ": unknown kind");
            if (kind == SHOCCS_SPHERE) {
                b.center = {q[1], q[2], q[3]};
                b.radius = q[4];
            } else {
                // make_rect (rect.cpp:8-19): the fluid normal is its sign
                b.kind = 1;
                b.axis = kind - SHOCCS_RECT_X;
                b.plane = q[1];
                b.c0[0] = q[2];
                b.c0[1] = q[3];
                b.c1[0] = q[4];
                b.c1[1] = q[5];
                b.fluid_normal = q[6] > 0 ? 1.0 : -1.0;
            }
        } else {
            const double* q = desc->spheres + 4 * s;
            b.center = {q[0], q[1], q[2]};
            b.radius = q[3];
        }
        shapes.push_back(b);
        p->obj.push_back((bcs::type)desc->shape_bc[s]);
    }
    for (int a = 0; a < 3; ++a) {
        p->grid[a].left = (bcs::type)desc->grid_bc[2 * a];
        p->grid[a].right = (bcs::type)desc->grid_bc[2 * a + 1];
    }
    p->m = std::make_unique<mesh>(int3{desc->n[0], desc->n[1], desc->n[2]},
                                  real3{desc->min[0], desc->min[1], desc->min[2]},
                                  real3{desc->max[0], desc->max[1], desc->max[2]}, shapes);
    // Inflow shapes resolve against the constant velocity of the linear wave
    const double* vel = desc->kind == SHX_SYS_WAVE_LINEAR ? pr : nullptr;
    if (!point_conditions(*p->m, p->obj, vel, p->pbc))
        return host_fail(SHX_EINVAL, "an inflow body needs the constant-velocity wave");
    if (desc->mesh_only) {
        *out = p.release();
        return SHX_OK;
    }
    for (int a = 0; a < 3; ++a) {
        p->d[a] = std::make_unique<derivative>(ctx, a, *p->m, *st, p->grid, p->pbc);
        if (p->d[a]->status() != SHX_OK) return host_fail(p->d[a]->status(), p->d[a]->error());
    }

    shx_system_desc sd;
    std::memset(&sd, 0, sizeof sd);
    sd.kind = desc->kind;
    const mesh& m = *p->m;
    for (int a = 0; a < 3; ++a) {
        sd.n[a] = m.extents()[a];
        sd.h[a] = m.h(a);
        sd.nR[a] = (int64_t)m.R(a).size();
        sd.d[a] = m.extents()[a] > 1 ? p->d[a]->device() : nullptr;
    }
    sd.x = m.coords(0).data();
    sd.y = m.coords(1).data();
    sd.z = m.coords(2).data();
    for (int f = 0; f < 6; ++f) sd.grid_bc[f] = desc->grid_bc[f];
    std::vector<double> rpos[3];
    std::vector<int32_t> rdir[3];
    for (int a = 0; a < 3; ++a) {
        for (size_t q = 0; q < m.R(a).size(); ++q) {
            const auto& r = m.R(a)[q];
            rpos[a].insert(rpos[a].end(), r.position.begin(), r.position.end());
            rdir[a].push_back(p->pbc[a][q] == bcs::type::Dirichlet);
        }
        rpos[a].push_back(0.0);
        rdir[a].push_back(0);
        sd.r_pos[a] = rpos[a].data();
        sd.r_dirichlet[a] = rdir[a].data();
    }
    sd.fluid = m.fluid().data();
    sd.nfluid = (int64_t)m.fluid().size();
    std::vector<double> mms;
    // valid() bounds: scalar_wave max_error (default 100, scalar_wave.cpp:137),
    // heat 1e6 (heat.cpp:54-58)
    if (desc->kind == SHX_SYS_WAVE_RADIAL && np > 4) p->max_error = pr[4];
    if (desc->kind == SHX_SYS_WAVE_LINEAR && np > 8) p->max_error = pr[8];
    if (desc->kind == SHX_SYS_HEAT) p->max_error = 1e6;
    if (desc->kind == SHX_SYS_WAVE_RADIAL) {
        for (int c = 0; c < 3; ++c) sd.center[c] = pr[c];
        sd.radius = pr[3];
    } else if (desc->kind == SHX_SYS_WAVE_LINEAR) {
        for (int c = 0; c < 3; ++c) {
            sd.vel[c] = pr[c];
            sd.x0[c] = pr[3 + c];
        }
        sd.width = pr[6];
        sd.amp = pr[7];
    } else {
        sd.diffusivity = pr[0];
        p->diffusivity = pr[0];
        sd.mms_kind = (int32_t)pr[1];
        sd.mms_dims = m.dims();
        if (sd.mms_kind == SHX_MMS_GAUSS) {
            if (np < 3 || np < 3 + 8 * (int)pr[2]) return SHX_EINVAL;
            sd.mms_n = (int32_t)pr[2];
            mms.assign(pr + 3, pr + 3 + 8 * sd.mms_n);
            sd.mms = mms.data();
        }
        if (sd.mms_kind == SHX_MMS_HOST) sd.mms_fns = desc->mms_fns;
    }
    // a NULL context builds the host structures only (no device, no system)
    int rc = ctx ? shx_system_create(ctx, &sd, &p->sys) : SHX_OK;
    if (rc != SHX_OK) return rc;
    *out = p.release();
    return SHX_OK;
}

int shoccs_problem_destroy(shoccs_problem* p)
{
    if (!p) return SHX_EINVAL;
    if (p->sys) shx_system_destroy(p->sys);
    delete p;
    return SHX_OK;
}

shx_system* shoccs_problem_system(shoccs_problem* p) { return p ? p->sys : nullptr; }

shx_derivative* shoccs_problem_derivative(shoccs_problem* p, int dir)
{
    if (!p || dir < 0 || dir > 2 || !p->d[dir]) return nullptr;
    return p->d[dir]->device();
}

int64_t shoccs_problem_size(const shoccs_problem* p, int which)
{
    if (!p) return -1;
    if (which == 0) return p->m->size();
    if (which >= 1 && which <= 3) return (int64_t)p->m->R(which - 1).size();
    if (which == 4) return (int64_t)p->m->fluid().size();
    return -1;
}

double shoccs_problem_dt(const shoccs_problem* p, double h_cfl, double p_cfl)
{
    const real3& h = p->m->h();
    const double hmin = std::min(h[0], std::min(h[1], h[2]));
    if (p->kind == SHX_SYS_HEAT) return p_cfl * hmin * hmin / (4 * p->diffusivity);
    return h_cfl * hmin;
}

int shoccs_euler_create(shoccs_problem* p, double gamma, shx_euler** out)
{
    if (!p || !out || !p->ctx) return SHX_EINVAL;
    const mesh& m = *p->m;
    const int3& n = m.extents();
    const int64_t st[3] = {(int64_t)n[1] * n[2], n[2], 1};
    std::vector<int64_t> fl[3];
    std::vector<double> nr[3];
    shx_euler_desc d;
    std::memset(&d, 0, sizeof d);
    d.gamma = gamma;
    for (int f = 0; f < 6; ++f) d.grid_bc[f] = (int32_t)(f % 2 ? p->grid[f / 2].right : p->grid[f / 2].left);
    for (int a = 0; a < 3; ++a) {
        // the fluid side of a cut point: solid + 1 along the ray, or - 1 when
        // the ray leaves the body (cut_discretization's closure columns,
        // src/operators/derivative.cpp:189-206)
        for (const auto& r : m.R(a)) {
            fl[a].push_back(m.ic(r.solid_coord) + (r.ray_outside ? -st[a] : st[a]));
            nr[a].insert(nr[a].end(), r.normal.begin(), r.normal.end());
        }
        fl[a].push_back(0);
        nr[a].push_back(0.0);
        d.wall_fluid[a] = fl[a].data();
        d.wall_normal[a] = nr[a].data();
    }
    shx_derivative* dv[3];
    for (int a = 0; a < 3; ++a) {
        if (!p->d[a]) return SHX_EINVAL;  // a mesh-only problem
        dv[a] = n[a] > 1 ? p->d[a]->device() : nullptr;
    }
    return shx_euler_create(p->ctx, dv, &d, out);
}

int shoccs_problem_geometry(const shoccs_problem* p, int32_t n[3], double lo[3], double hi[3])
{
    if (!p || !n || !lo || !hi) return SHX_EINVAL;
    for (int a = 0; a < 3; ++a) {
        n[a] = p->m->extents()[a];
        lo[a] = p->m->lo()[a];
        hi[a] = p->m->hi()[a];
    }
    return SHX_OK;
}

int shoccs_error_field(const shoccs_problem* p, const double* u, const double* sol, double* err)
{
    if (!p || !u || !sol || !err) return SHX_EINVAL;
    const mesh& m = *p->m;
    const int64_t N = m.size();
    int64_t total = N;
    for (int a = 0; a < 3; ++a) total += (int64_t)m.R(a).size();
    std::fill(err, err + total, 0.0);
    for (int64_t i : m.fluid()) err[i] = std::abs(u[i] - sol[i]);
    int64_t off = N;
    for (int a = 0; a < 3; ++a) {
        const auto& R = m.R(a);
        for (size_t i = 0; i < R.size(); ++i)
            if (p->pbc[a][i] != bcs::type::Dirichlet)
                err[off + (int64_t)i] = std::abs(u[off + (int64_t)i] - sol[off + (int64_t)i]);
        off += (int64_t)R.size();
    }
    // Dirichlet grid faces carry no error
    const int3& n = m.extents();
    for (int a = 0; a < 3; ++a) {
        const bcs::type side[2] = {p->grid[a].left, p->grid[a].right};
        for (int s = 0; s < 2; ++s) {
            if (side[s] != bcs::type::Dirichlet) continue;
            const int plane = s == 0 ? 0 : n[a] - 1;
            for (int i = 0; i < n[0]; ++i)
                for (int j = 0; j < n[1]; ++j)
                    for (int k = 0; k < n[2]; ++k) {
                        const int q[3] = {i, j, k};
                        if (q[a] == plane) err[m.ic({i, j, k})] = 0.0;
                    }
        }
    }
    return SHX_OK;
}

int shoccs_mesh_r(const shoccs_problem* p, int dir, double* psi, double* pos, int32_t* solid,
                  int32_t* ray_outside, int32_t* shape_id)
{
    if (!p || dir < 0 || dir > 2) return SHX_EINVAL;
    const auto& R = p->m->R(dir);
    for (size_t i = 0; i < R.size(); ++i) {
        if (psi) psi[i] = R[i].psi;
        for (int c = 0; c < 3; ++c) {
            if (pos) pos[3 * i + c] = R[i].position[c];
            if (solid) solid[3 * i + c] = R[i].solid_coord[c];
        }
        if (ray_outside) ray_outside[i] = R[i].ray_outside;
        if (shape_id) shape_id[i] = R[i].shape_id;
    }
    return SHX_OK;
}

int64_t shoccs_mesh_nlines(const shoccs_problem* p, int dir)
{
    return p && dir >= 0 && dir < 3 ? (int64_t)p->m->lines(dir).size() : -1;
}

int shoccs_mesh_lines(const shoccs_problem* p, int dir, int64_t* out)
{
    if (!p || dir < 0 || dir > 2 || !out) return SHX_EINVAL;
    const int S = slow_axis(dir), F = fast_axis(dir);
    int64_t* o = out;
    for (const auto& l : p->m->lines(dir)) {
        o[0] = l.start.mesh_coordinate[dir];
        o[1] = l.end.mesh_coordinate[dir];
        o[2] = l.start.object ? l.start.object->object_coordinate : -1;
        o[3] = l.end.object ? l.end.object->object_coordinate : -1;
        o[4] = l.start.mesh_coordinate[S];
        o[5] = l.start.mesh_coordinate[F];
        o[6] = l.stride;
        o += 7;
    }
    return SHX_OK;
}

int shoccs_mesh_fluid(const shoccs_problem* p, int64_t* out)
{
    if (!p || !out) return SHX_EINVAL;
    const auto& f = p->m->fluid();
    std::memcpy(out, f.data(), f.size() * sizeof(int64_t));
    return SHX_OK;
}

int64_t shoccs_deriv_nblocks(const shoccs_problem* p, int dir)
{
    return p && dir >= 0 && dir < 3 && p->d[dir] ? (int64_t)p->d[dir]->blocks().size() : -1;
}

int shoccs_deriv_blocks(const shoccs_problem* p, int dir, shx_line_block* out)
{
    if (!p || dir < 0 || dir > 2 || !out || !p->d[dir]) return SHX_EINVAL;
    const auto& b = p->d[dir]->blocks();
    std::memcpy(out, b.data(), b.size() * sizeof(shx_line_block));
    return SHX_OK;
}

int64_t shoccs_deriv_npool(const shoccs_problem* p, int dir)
{
    return p && dir >= 0 && dir < 3 && p->d[dir] ? (int64_t)p->d[dir]->pool().size() : -1;
}

int shoccs_deriv_pool(const shoccs_problem* p, int dir, double* out)
{
    if (!p || dir < 0 || dir > 2 || !out || !p->d[dir]) return SHX_EINVAL;
    const auto& v = p->d[dir]->pool();
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return SHX_OK;
}

int64_t shoccs_deriv_ncut(const shoccs_problem* p, int dir, int64_t* nnz)
{
    if (!p || dir < 0 || dir > 2 || !p->d[dir]) return -1;
    if (nnz) *nnz = (int64_t)p->d[dir]->rw().size();
    return (int64_t)p->d[dir]->rrow().size();
}

int shoccs_deriv_cut(const shoccs_problem* p, int dir, int32_t* rset, int64_t* rrow, int64_t* rptr,
                     int64_t* rcol, double* rw)
{
    if (!p || dir < 0 || dir > 2 || !rset || !rrow || !rptr || !rcol || !rw || !p->d[dir]) return SHX_EINVAL;
    const derivative& d = *p->d[dir];
    std::memcpy(rset, d.rset().data(), d.rset().size() * sizeof(int32_t));
    std::memcpy(rrow, d.rrow().data(), d.rrow().size() * sizeof(int64_t));
    std::memcpy(rptr, d.rptr().data(), d.rptr().size() * sizeof(int64_t));
    std::memcpy(rcol, d.rcol().data(), d.rcol().size() * sizeof(int64_t));
    std::memcpy(rw, d.rw().data(), d.rw().size() * sizeof(double));
    return SHX_OK;
}

int shoccs_deriv_interior(const shoccs_problem* p, int dir, double* out9)
{
    if (!p || dir < 0 || dir > 2 || !out9 || !p->d[dir]) return SHX_EINVAL;
    const auto& v = p->d[dir]->interior();
    for (size_t j = 0; j < 9; ++j) out9[j] = j < v.size() ? v[j] : 0.0;
    return (int)v.size();
}

int shoccs_field_alloc(shoccs_problem* p, shx_field* f)
{
    if (!p || !f) return SHX_EINVAL;
    std::memset(f, 0, sizeof *f);
    int rc = shx_alloc(p->ctx, sizeof(double) * (size_t)p->m->size(), (void**)&f->D);
    for (int a = 0; a < 3 && rc == SHX_OK; ++a)
        rc = shx_alloc(p->ctx, sizeof(double) * (p->m->R(a).size() + 1), (void**)&f->R[a]);
    if (rc == SHX_OK) rc = shx_fill(p->ctx, f->D, 0.0, p->m->size());
    for (int a = 0; a < 3 && rc == SHX_OK; ++a) rc = shx_fill(p->ctx, f->R[a], 0.0, (int64_t)p->m->R(a).size());
    if (rc == SHX_OK) rc = shx_sync(p->ctx);
    return rc;
}

int shoccs_field_free(shoccs_problem* p, shx_field* f)
{
    if (!p || !f) return SHX_EINVAL;
    shx_free(p->ctx, f->D);
    for (int a = 0; a < 3; ++a) shx_free(p->ctx, f->R[a]);
    std::memset(f, 0, sizeof *f);
    return SHX_OK;
}

int shoccs_field_upload(shoccs_problem* p, shx_field* f, const double* flat)
{
    if (!p || !f || !flat) return SHX_EINVAL;
    int rc = shx_h2d(p->ctx, f->D, flat, sizeof(double) * (size_t)p->m->size());
    size_t off = (size_t)p->m->size();
    for (int a = 0; a < 3 && rc == SHX_OK; ++a) {
        rc = shx_h2d(p->ctx, f->R[a], flat + off, sizeof(double) * p->m->R(a).size());
        off += p->m->R(a).size();
    }
    return rc;
}

int shoccs_field_download(shoccs_problem* p, const shx_field* f, double* flat)
{
    if (!p || !f || !flat) return SHX_EINVAL;
    int rc = shx_d2h(p->ctx, flat, f->D, sizeof(double) * (size_t)p->m->size());
    size_t off = (size_t)p->m->size();
    for (int a = 0; a < 3 && rc == SHX_OK; ++a) {
        rc = shx_d2h(p->ctx, flat + off, f->R[a], sizeof(double) * p->m->R(a).size());
        off += p->m->R(a).size();
    }
    return rc;
}

namespace {

int copy_field(shx_ctx* ctx, const shoccs_problem* p, shx_field* dst, const shx_field* src)
{
    int rc = shx_copy(ctx, dst->D, src->D, shoccs_problem_size(p, 0));
    for (int a = 0; a < 3 && rc == SHX_OK; ++a) {
        const int64_t n = shoccs_problem_size(p, a + 1);
        if (n > 0) rc = shx_copy(ctx, dst->R[a], src->R[a], n);
    }
    return rc;
}

}  // namespace

int shoccs_run(shoccs_problem* p, const shoccs_run_desc* d, shx_field* u, shoccs_run_result* r)
{
    if (!p || !d || !u || !r || !p->sys || !p->ctx) return SHX_EINVAL;
    if (d->integrator < SHOCCS_RK4 || d->integrator > SHOCCS_EULER) return SHX_EINVAL;
    std::memset(r, 0, sizeof *r);
    r->linf = std::numeric_limits<double>::quiet_NaN();
    int max_step = d->max_step;
    const double max_time = d->max_time;
    if (max_step == INT32_MAX && max_time == DBL_MAX) max_step = 0;
    shx_system* sys = p->sys;
    shx_ctx* ctx = p->ctx;
    const bool fused = d->integrator == SHOCCS_RK4_FUSED;

    // u1 (the step output), rk and srhs registers of the reference schedule
    shx_field u1{}, rk{}, srhs{};
    int rc = shoccs_field_alloc(p, &u1);
    if (rc == SHX_OK && d->integrator == SHOCCS_RK4) rc = shoccs_field_alloc(p, &rk);
    if (rc == SHX_OK && !fused) rc = shoccs_field_alloc(p, &srhs);

    double stats[11] = {0};
    double t = 0.0;
    int step = 0;
    auto record = [&]() {
        if (d->stats && r->nstats < d->stats_cap) {
            std::memcpy(d->stats + 11 * r->nstats, stats, sizeof stats);
            ++r->nstats;
        }
    };
    auto running = [&]() { return 0.0 <= t && t < max_time && 0 <= step && step < max_step; };
    auto valid = [&]() { return std::isfinite(stats[0]) && std::abs(stats[0]) <= p->max_error; };

    // simulation_cycle.cpp:59-64: initialize u0, u1 <- u0, update u0's
    // boundary, statistics of u1
    if (rc == SHX_OK) rc = shx_system_initialize(sys, u, 0.0);
    if (rc == SHX_OK) rc = copy_field(ctx, p, &u1, u);
    if (rc == SHX_OK) rc = shx_system_update_boundary(sys, u, 0.0);
    if (rc == SHX_OK) rc = shx_system_stats(sys, &u1, 0.0, stats);
    if (rc == SHX_OK) record();
    // system::write (scalar_system_utils.hpp:224-301) through the field
    // writer: U and its Error against the exact solution
    std::vector<double> hu, hs, he;
    shx_field ex{};
    auto write = [&](const shx_field* f, int32_t k, double time, double dt) {
        if (!d->io) return SHX_OK;
        int r2 = SHX_OK;
        if (!ex.D) {
            r2 = shoccs_field_alloc(p, &ex);
            int64_t total = shoccs_problem_size(p, 0);
            for (int a = 1; a <= 3; ++a) total += shoccs_problem_size(p, a);
            hu.resize((size_t)total);
            hs.resize((size_t)total);
            he.resize((size_t)total);
        }
        if (r2 == SHX_OK) r2 = shx_system_exact(sys, &ex, time);
        if (r2 == SHX_OK) r2 = shoccs_field_download(p, f, hu.data());
        if (r2 == SHX_OK) r2 = shoccs_field_download(p, &ex, hs.data());
        if (r2 == SHX_OK) r2 = shoccs_error_field(p, hu.data(), hs.data(), he.data());
        const char* names[2] = {"U", "Error"};
        const double* fields[2] = {hu.data(), he.data()};
        if (r2 == SHX_OK) r2 = shoccs_io_write(d->io, 2, names, fields, k, time, dt, nullptr);
        return r2;
    };
    if (rc == SHX_OK) rc = write(u, 0, 0.0, 0.0);
    bool dt_ok = true;
    while (rc == SHX_OK && running() && valid()) {
        const double dt = shoccs_problem_dt(p, d->h_cfl, d->p_cfl);
        if (!(dt >= d->min_dt)) {  // step_controller::check_timestep_size
            dt_ok = false;
            break;
        }
        if (d->integrator == SHOCCS_RK4) {
            rc = shx_rk4_step(sys, u, &u1, &rk, &srhs, t, dt);
        } else if (d->integrator == SHOCCS_EULER) {
            rc = shx_euler_step(sys, u, &u1, &srhs, t, dt);
        } else {
            rc = shx_rk4_fused_steps_async(sys, u, t, dt, 1);
            if (rc == SHX_OK) rc = shx_sync(ctx);
        }
        if (rc != SHX_OK) break;
        t += dt;  // controller.advance
        step += 1;
        rc = shx_system_stats(sys, fused ? u : &u1, t, stats);
        if (rc == SHX_OK) record();
        if (rc == SHX_OK) rc = write(fused ? u : &u1, step, t, dt);
        if (rc == SHX_OK && !fused) rc = copy_field(ctx, p, u, &u1);  // u0 <- u1
    }
    if (rc == SHX_OK) rc = shx_sync(ctx);
    shoccs_field_free(p, &u1);
    if (ex.D) shoccs_field_free(p, &ex);
    if (rk.D) shoccs_field_free(p, &rk);
    if (srhs.D) shoccs_field_free(p, &srhs);
    if (rc != SHX_OK) return rc;
    r->time = t;
    r->step = step;
    r->completed = dt_ok && !running();
    if (r->completed) r->linf = stats[0];
    return SHX_OK;
}

}  // extern "C"
