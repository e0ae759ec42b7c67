// mesh.hpp -- Cartesian mesh with embedded spheres and rectangles, found by ray casting.
//
// Mirrors ccs::mesh (src/mesh/mesh.hpp:15-119) and ccs::object_geometry
// (src/mesh/object_geometry.hpp): coordinates, cut points R(dir) ordered by
// (slow, fast) then along dir, line segments per direction, and the fluid
// gather.  Runs once on the host; every value that decides a cut point is
// computed in the reference's order so that masks, line extents and psi come
// out identical.
#pragma once

#include <array>
#include <cstdint>
#include <optional>
#include <vector>

namespace ccs {

using real = double;
using real3 = std::array<real, 3>;
using int3 = std::array<int, 3>;

namespace bcs {
// src/operators/boundaries.hpp:13-20.  Inflow (no reference counterpart) is a
// shape condition of the constant-velocity wave that each cut point resolves
// to Dirichlet or Floating (see point_conditions); operators only ever see
// the resolved conditions.
enum class type : int { Dirichlet = 0, Floating = 1, Neumann = 2, Inflow = 3 };
struct Line {
    type left = type::Floating;
    type right = type::Floating;
};
using Grid = std::array<Line, 3>;
using Object = std::vector<type>;  // one condition per shape
using Points = std::array<std::vector<type>, 3>;  // one per cut point of R(x), R(y), R(z)
}  // namespace bcs

// index::dir<I>::slow / fast (src/indexing.hpp:21-37)
constexpr int slow_axis(int dir) { return dir == 0 ? 1 : 0; }
constexpr int fast_axis(int dir) { return dir == 2 ? 1 : 2; }

// A body: a sphere (src/mesh/sphere.cpp:9-45), or a rectangle in the plane
// normal to `axis` at coordinate `plane`, bounded by c0 - c1 in that plane's
// (slow, fast) axes, with the fluid on the side of fluid_normal (+1 or -1;
// src/mesh/rect.hpp:11-44, rect.cpp:8-19).
struct body_shape {
    int kind = 0;  // 0 sphere, 1 rectangle
    real3 center{};
    real radius = 0;
    int axis = 0;
    real plane = 0;
    real c0[2] = {0, 0}, c1[2] = {0, 0};
    real fluid_normal = 1;
};

// src/mesh/mesh_types.hpp:30-38
struct mesh_object_info {
    real psi;
    real3 position;
    real3 normal;
    bool ray_outside;
    int3 solid_coord;
    int shape_id;
};

struct object_boundary {
    int64_t object_coordinate;
    int64_t objectID;
    real psi;
};

struct boundary {
    int3 mesh_coordinate;
    std::optional<object_boundary> object;
};

struct line {
    int64_t stride;
    boundary start;
    boundary end;
};

class mesh;
// The condition of every cut point: its shape's condition, with Inflow
// resolved against the velocity vel (nullptr: Inflow is refused) to
// Dirichlet where vel . n > 0 at the cut point (n the body's outward normal:
// the flow leaves the body into the fluid) and Floating elsewhere -- the
// reference's pattern of a Dirichlet body upstream and a floating one
// downstream (scalar_wave.lua:10-22), applied point by point.  Returns false
// when an Inflow shape has no velocity to resolve it.
bool point_conditions(const mesh& m, const bcs::Object& obj, const double* vel, bcs::Points& out);

class mesh
{
public:
    mesh() = default;
    mesh(const int3& n, const real3& min, const real3& max, const std::vector<body_shape>& shapes);

    const int3& extents() const { return n_; }
    int64_t size() const { return (int64_t)n_[0] * n_[1] * n_[2]; }
    int dims() const { return dims_; }
    real h(int i) const { return h_[i]; }
    const real3& h() const { return h_; }
    const std::vector<real>& coords(int i) const { return xyz_[i]; }
    const std::vector<mesh_object_info>& R(int dir) const { return R_[dir]; }
    const std::vector<line>& lines(int dir) const { return lines_[dir]; }
    const std::vector<int64_t>& fluid() const { return fluid_; }
    const real3& lo() const { return min_; }
    const real3& hi() const { return max_; }
    const std::vector<body_shape>& shapes() const { return shapes_; }
    int64_t stride(int dir) const;
    int64_t ic(const int3& ijk) const
    {
        return (int64_t)ijk[0] * n_[1] * n_[2] + (int64_t)ijk[1] * n_[2] + ijk[2];
    }
    // mesh::dirichlet_line (src/mesh/mesh.cpp:154-168)
    bool dirichlet_line(const int3& start, int dir, const bcs::Grid& grid) const;
    // mesh::interp_line (src/mesh/mesh.cpp:170-207): the line along dir
    // through pt, bounded by the mesh ends or by the nearest cut points of
    // R(dir) on either side of pt
    line interp_line(int dir, const int3& pt) const;

private:
    int3 n_{1, 1, 1};
    real3 min_{}, max_{}, h_{};
    int dims_ = 0;
    std::array<std::vector<real>, 3> xyz_;
    std::vector<body_shape> shapes_;
    std::array<std::vector<mesh_object_info>, 3> R_;
    std::array<std::vector<line>, 3> lines_;
    std::vector<int64_t> fluid_;

    void cast_rays(int dir);
    void build_lines(int dir);
    void build_fluid();
};

}  // namespace ccs
