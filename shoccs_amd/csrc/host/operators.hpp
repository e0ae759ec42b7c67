// operators.hpp -- derivative, gradient and laplacian on one GPU.
//
// ccs::derivative assembles one direction's operator on the host, as the
// reference's domain_discretization (src/operators/derivative.cpp:305-449),
// but emits it directly in the device layout: one shx_line_block per line
// segment (positions instead of flat offsets), closure rows and body B
// weights in a de-duplicated coefficient pool, Neumann rows as triplets.
// Application goes through the C ABI of shoccs_hip.h.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "mesh.hpp"
#include "shoccs_hip.h"
#include "stencils.hpp"

namespace ccs {

class derivative
{
public:
    // pbc: the condition of every cut point (point_conditions)
    derivative(shx_ctx* ctx, int dir, const mesh& m, const stencil& st, const bcs::Grid& grid,
               const bcs::Points& pbc);
    derivative(shx_ctx* ctx, int dir, const mesh& m, const stencil& st, const bcs::Grid& grid,
               const bcs::Object& obj);
    ~derivative();
    derivative(const derivative&) = delete;
    derivative& operator=(const derivative&) = delete;

    int status() const { return status_; }
    const std::string& error() const { return error_; }
    int dir() const { return dir_; }
    shx_derivative* device() const { return dev_; }
    const std::vector<shx_line_block>& blocks() const { return blocks_; }
    const std::vector<double>& pool() const { return pool_; }
    const std::vector<double>& interior() const { return interior_; }
    // R-space rows (the cut-point rows of non-Dirichlet bodies): row k is row
    // rrow[k] of R(rset[k]) with entries rptr[k]..rptr[k+1]; rcol >= 0 names a
    // point of D, -1 - c point c of R(rset[k])
    const std::vector<int32_t>& rset() const { return rset_; }
    const std::vector<int64_t>& rrow() const { return rrow_; }
    const std::vector<int64_t>& rptr() const { return rptr_; }
    const std::vector<int64_t>& rcol() const { return rcol_; }
    const std::vector<double>& rw() const { return rw_; }

    // derivative::operator()(u, du, op) (src/operators/derivative.cpp:502-509)
    int operator()(const shx_field& u, shx_field& du, bool plus_eq = false) const;

private:
    int dir_ = 0;
    int status_ = SHX_OK;
    std::string error_;
    shx_derivative* dev_ = nullptr;
    std::vector<shx_line_block> blocks_;
    std::vector<double> pool_;
    std::vector<double> interior_;
    std::vector<int64_t> neu_row_, neu_col_;
    std::vector<double> neu_w_;
    std::vector<int32_t> rset_;
    std::vector<int64_t> rrow_, rptr_, rcol_;
    std::vector<double> rw_;
};

// ccs::gradient / ccs::laplacian (src/operators/gradient.cpp:13-46,
// src/operators/laplacian.cpp:14-58): three derivatives.
class gradient
{
public:
    gradient(shx_ctx* ctx, const mesh& m, const stencil& st, const bcs::Grid& grid,
             const bcs::Object& obj);
    int operator()(const shx_field& u, shx_field& dux, shx_field& duy, shx_field& duz) const;
    const derivative* d(int i) const { return d_[i].get(); }
    int status() const;

private:
    std::unique_ptr<derivative> d_[3];
};

class laplacian
{
public:
    laplacian(shx_ctx* ctx, const mesh& m, const stencil& st, const bcs::Grid& grid,
              const bcs::Object& obj);
    int operator()(const shx_field& u, shx_field& du) const;
    const derivative* d(int i) const { return d_[i].get(); }
    int status() const;

private:
    std::unique_ptr<derivative> d_[3];
};

}  // namespace ccs
