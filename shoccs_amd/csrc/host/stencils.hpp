// stencils.hpp -- the interior stencil and numerical boundary scheme of one
// discretisation, as the type-erased ccs::stencil of the reference
// (src/stencils/stencil.hpp:22-300, src/stencils/stencil.cpp:8-106).
#pragma once

#include <array>
#include <optional>

#include "mesh.hpp"

namespace ccs::stencils {

// stencil.hpp:22-27: interior order 2p with 2p+1 points; r x t closure; extras
struct info {
    int p = 0;
    int r = 0;
    int t = 0;
    int nextra = 0;
};

// stencil.hpp:29 interp_info: p interior interpolation points, t wall points
struct interp_info {
    int p = 0;
    int t = 0;
};

// scheme{order, type}: order 2 "E2"/"E4" and order 1 "E4"/"E4u"/"E2-poly"
enum class scheme : int { E2_2 = 0, E4_2 = 1, E4_1 = 2, E4u_1 = 3, E2poly_1 = 4 };

// the most free parameters a scheme takes (E2-poly: floating 6, Dirichlet 3,
// interpolant 4, in that order)
constexpr int kMaxAlpha = 13;

class stencil
{
public:
    stencil() = default;
    stencil(scheme k, std::array<double, kMaxAlpha> alpha) : kind_(k), alpha_(alpha) {}

    // returns nullopt for an unknown scheme or an E4_1 alpha[1] below 197/288
    static std::optional<stencil> make(int kind, const double* alpha, int nalpha);

    info query(bcs::type b) const;
    info query_max() const;
    void interior(double h, double* c) const;
    // closure of r x t coefficients (row-major) for a wall at distance psi*h;
    // Neumann closures also fill `extra`
    void nbs(double h, bcs::type b, double psi, bool right, double* c, double* extra) const;
    // interpolation weights written to c, their count returned; schemes
    // without interpolants leave c as it is and return cap
    interp_info query_interp() const;
    int interp_interior(double y, double* c, int cap) const;
    int interp_wall(int i, double y, double psi, double* c, bool right, int cap) const;
    scheme kind() const { return kind_; }

private:
    scheme kind_ = scheme::E4_2;
    std::array<double, kMaxAlpha> alpha_{};
};

}  // namespace ccs::stencils

namespace ccs {
using stencils::stencil;
}
