// mesh.cpp -- Cartesian mesh, ray casting and line segments.
//
//   cartesian ctor / linear_distribute  src/mesh/cartesian.cpp:13-44,
//                                       src/fields/lazy_views.hpp:563-569
//   sphere::hit / normal                src/mesh/sphere.cpp:15-45
//   rect<I>::hit / normal               src/mesh/rect.hpp:21-43
//   dot (unary right fold)              src/real3_operators.hpp:62-72
//   ray casting, snap and psi clamp     src/mesh/object_geometry.cpp:17-116
//   line segments                       src/mesh/mesh.cpp:29-87
//   fluid gather                        src/mesh/mesh.cpp:89-111,134-137
#include "mesh.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace ccs {

namespace {

real dot(const real3& a, const real3& b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }

real3 sub(const real3& a, const real3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }

struct hit_info {
    real t;
    real3 position;
    bool ray_outside;
    int shape_id;
};

std::optional<hit_info> sphere_hit(const body_shape& s, int id, const real3& origin,
                                   const real3& direction, real t_min, real t_max)
{
    const real3 oc = sub(origin, s.center);
    const real a = dot(direction, direction);
    const real b = dot(oc, direction);
    const real c = dot(oc, oc) - s.radius * s.radius;
    const real disc = b * b - a * c;
    if (disc > 0) {
        const real sqr = std::sqrt(disc);
        for (real t : {(-b - sqr) / a, (-b + sqr) / a}) {
            if (t > t_min && t < t_max) {
                real3 p;
                for (int q = 0; q < 3; ++q) p[q] = origin[q] + t * direction[q];
                return hit_info{t, p, dot(direction, sub(p, s.center)) < 0, id};
            }
        }
    }
    return std::nullopt;
}

// rect<I>::hit (src/mesh/rect.hpp:21-36): bounds and t_min, t_max inclusive.
// A line lying in the plane (t = 0/0) takes no hit; the reference computes
// with a NaN t there.
std::optional<hit_info> rect_hit(const body_shape& s, int id, const real3& origin, const real3& direction,
                                 real t_min, real t_max)
{
    const int I = s.axis, S = slow_axis(I), F = fast_axis(I);
    const real t = (s.plane - origin[I]) / direction[I];
    if (t < t_min || t > t_max || std::isnan(t)) return std::nullopt;
    real3 p;
    for (int q = 0; q < 3; ++q) p[q] = origin[q] + t * direction[q];
    if (p[S] < s.c0[0] || p[S] > s.c1[0] || p[F] < s.c0[1] || p[F] > s.c1[1]) return std::nullopt;
    return hit_info{t, p, s.fluid_normal * direction[I] < 0, id};
}

real3 shape_normal(const body_shape& s, const real3& pos)
{
    if (s.kind == 1) {  // rect<I>::normal (rect.hpp:38-43)
        real3 n{0, 0, 0};
        n[s.axis] = s.fluid_normal;
        return n;
    }
    const real3 r = sub(pos, s.center);
    const real d = std::sqrt(dot(r, r));
    return {r[0] / d, r[1] / d, r[2] / d};
}

}  // namespace

mesh::mesh(const int3& n, const real3& min, const real3& max, const std::vector<body_shape>& shapes)
    : shapes_(shapes)
{
    for (int i = 0; i < 3; ++i) {
        n_[i] = n[i] > 0 ? n[i] : 1;
        min_[i] = min[i];
        max_[i] = max[i];
        h_[i] = (n_[i] - 1) ? (max_[i] - min_[i]) / (n_[i] - 1) : DBL_MAX;
        if (n_[i] != 1) ++dims_;
        xyz_[i].resize(n_[i]);
        for (int k = 0; k < n_[i]; ++k)
            xyz_[i][k] = n_[i] > 1 ? min_[i] + k * (max_[i] - min_[i]) / (n_[i] - 1) : min_[i];
    }
    for (int d = 0; d < 3; ++d) cast_rays(d);
    for (int d = 0; d < 3; ++d) build_lines(d);
    build_fluid();
}

int64_t mesh::stride(int dir) const
{
    int64_t s = 1;
    for (int i = dir + 1; i < 3; ++i) s *= n_[i];
    return s;
}

void mesh::cast_rays(int I)
{
    const int S = slow_axis(I), F = fast_axis(I);
    auto& out = R_[I];
    real3 origin{};
    int3 coord{};
    for (int s = 0; s < n_[S]; ++s) {
        origin[S] = min_[S] + s * h_[S];
        coord[S] = s;
        for (int f = 0; f < n_[F]; ++f) {
            origin[F] = min_[F] + f * h_[F];
            coord[F] = f;
            real t_min = 0;
            real t_max = max_[I] - min_[I];
            origin[I] = min_[I];
            real3 direction{};
            direction[I] = 1.0;
            while (true) {
                // closest hit over all shapes (object_geometry.cpp:17-30)
                std::optional<hit_info> hit;
                real tm = t_max;
                for (int id = 0; id < (int)shapes_.size(); ++id)
                    if (auto h = shapes_[id].kind == 0 ? sphere_hit(shapes_[id], id, origin, direction, t_min, tm)
                                                       : rect_hit(shapes_[id], id, origin, direction, t_min, tm)) {
                        hit = h;
                        tm = h->t;
                    }
                if (!hit) break;
                const real t_over_h = hit->t / h_[I];
                int i_cell = (int)std::round(t_over_h);
                constexpr real snap_tol = 1e-12;
                if (std::abs(t_over_h - i_cell) > snap_tol) i_cell = (int)t_over_h;
                coord[I] = i_cell + hit->ray_outside;
                const int off = 1 - 2 * hit->ray_outside;
                const real fluid_pos = min_[I] + h_[I] * (coord[I] + off);
                real psi = off * (fluid_pos - hit->position[I]) / h_[I];
                psi = std::clamp(psi, snap_tol, 1.0 - snap_tol);
                out.push_back(mesh_object_info{psi, hit->position,
                                               shape_normal(shapes_[hit->shape_id], hit->position),
                                               hit->ray_outside, coord, hit->shape_id});
                t_min = std::nextafter(hit->t, t_max);
            }
        }
    }
}

void mesh::build_lines(int I)
{
    if (n_[I] == 1) return;
    const int S = slow_axis(I), F = fast_axis(I);
    const auto& r = R_[I];
    auto& v = lines_[I];
    v.reserve((size_t)n_[S] * n_[F] + r.size());
    const int64_t st = stride(I);
    size_t first = 0;
    for (int s = 0; s < n_[S]; ++s) {
        for (int f = 0; f < n_[F]; ++f) {
            int3 left{}, right{};
            left[S] = right[S] = s;
            left[F] = right[F] = f;
            left[I] = 0;
            right[I] = n_[I] - 1;
            std::optional<boundary> lb = boundary{left, std::nullopt};
            while (first < r.size() && r[first].solid_coord[S] == s && r[first].solid_coord[F] == f) {
                const boundary ob{r[first].solid_coord,
                                  object_boundary{(int64_t)first, r[first].shape_id, r[first].psi}};
                if (r[first].ray_outside) {
                    v.push_back(line{st, *lb, ob});
                    lb.reset();
                } else {
                    lb = ob;
                }
                ++first;
            }
            if (lb) v.push_back(line{st, *lb, boundary{right, std::nullopt}});
        }
    }
}

void mesh::build_fluid()
{
    const int i = n_[2] > 1 ? 2 : n_[1] > 1 ? 1 : 0;
    std::vector<std::pair<int64_t, int64_t>> slices;
    for (const auto& l : lines_[i]) {
        const int64_t i0 = l.start.object ? ic(l.start.mesh_coordinate) + 1 : ic(l.start.mesh_coordinate);
        const int64_t i1 = l.end.object ? ic(l.end.mesh_coordinate) : ic(l.end.mesh_coordinate) + 1;
        if (!slices.empty() && slices.back().second == i0) slices.back().second = i1;
        else slices.emplace_back(i0, i1);
    }
    for (const auto& [a, b] : slices)
        for (int64_t q = a; q < b; ++q) fluid_.push_back(q);
}

bool mesh::dirichlet_line(const int3& start, int dir, const bcs::Grid& grid) const
{
    bool result = false;
    for (int i = 0; i < 3; ++i) {
        if (i == dir) continue;
        const bool f = (grid[i].left == bcs::type::Dirichlet && start[i] == 0) ||
                       (grid[i].right == bcs::type::Dirichlet && start[i] == n_[i] - 1);
        result = result || f;
    }
    return result;
}

line mesh::interp_line(int dir, const int3& pt) const
{
    line out{stride(dir), boundary{pt, std::nullopt}, boundary{pt, std::nullopt}};
    out.start.mesh_coordinate[dir] = 0;
    out.end.mesh_coordinate[dir] = n_[dir] - 1;
    const int S = slow_axis(dir), F = fast_axis(dir);
    // R(dir) is ordered by (slow, fast): stop once past pt's slow coordinate
    const auto& cuts = R_[dir];
    for (int64_t i = 0; i < (int64_t)cuts.size(); ++i) {
        const mesh_object_info& o = cuts[i];
        if (o.solid_coord[S] > pt[S]) break;
        if (o.solid_coord[S] != pt[S] || o.solid_coord[F] != pt[F]) continue;
        boundary* b = nullptr;
        if (o.normal[dir] >= 0.0 && o.solid_coord[dir] <= pt[dir]) {
            // >=: a solid point on the mesh end still bounds the line
            if (o.solid_coord[dir] >= out.start.mesh_coordinate[dir]) b = &out.start;
        } else if (o.normal[dir] < 0.0 && o.solid_coord[dir] >= pt[dir]) {
            if (o.solid_coord[dir] <= out.end.mesh_coordinate[dir]) b = &out.end;
        }
        if (b) *b = boundary{o.solid_coord, object_boundary{i, o.shape_id, o.psi}};
    }
    return out;
}

bool point_conditions(const mesh& m, const bcs::Object& obj, const double* vel, bcs::Points& out)
{
    for (int d = 0; d < 3; ++d) {
        out[d].clear();
        for (const mesh_object_info& r : m.R(d)) {
            bcs::type bc = obj[r.shape_id];
            if (bc == bcs::type::Inflow) {
                if (!vel) return false;
                const double an = vel[0] * r.normal[0] + vel[1] * r.normal[1] + vel[2] * r.normal[2];
                bc = an > 0.0 ? bcs::type::Dirichlet : bcs::type::Floating;
            }
            out[d].push_back(bc);
        }
    }
    return true;
}

}  // namespace ccs
