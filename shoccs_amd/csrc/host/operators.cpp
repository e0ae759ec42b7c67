// operators.cpp -- host assembly of the derivative operators into the device
// segment table, and their application through shoccs_hip.h.
//
//   submatrix_size / domain_discretization  src/operators/derivative.cpp:259-449
//   inner_block placement of L | C | R      src/matrices/inner_block.cpp:17-40
//   derivative ctor                         src/operators/derivative.cpp:452-473
//   gradient / laplacian                    src/operators/gradient.cpp:34-46,
//                                           src/operators/laplacian.cpp:37-58
#include "operators.hpp"

#include <cstring>
#include <string>
#include <unordered_map>

namespace ccs {

using stencils::info;

namespace {

// Appends coefficient runs to the pool, sharing identical runs (the grid-wall
// closures of every line, psi-independent body closures).
class pool_builder
{
public:
    explicit pool_builder(std::vector<double>& pool) : pool_(pool) {}
    int32_t add(const double* v, int n)
    {
        std::string key(reinterpret_cast<const char*>(v), sizeof(double) * (size_t)n);
        auto it = index_.find(key);
        if (it != index_.end()) return it->second;
        const int32_t off = (int32_t)pool_.size();
        pool_.insert(pool_.end(), v, v + n);
        index_.emplace(std::move(key), off);
        return off;
    }

private:
    std::vector<double>& pool_;
    std::unordered_map<std::string, int32_t> index_;
};

}  // namespace

derivative::derivative(shx_ctx* ctx, int dir, const mesh& m, const stencil& st,
                       const bcs::Grid& grid, const bcs::Object& obj)
    : dir_(dir)
{
    const int3& n = m.extents();
    const info mx = st.query_max();
    interior_.assign(2 * mx.p + 1, 0.0);
    const double h = m.h(dir);
    if (n[dir] >= 2) {
        st.interior(h, interior_.data());
        std::vector<double> left(mx.r * mx.t + 8), right(mx.r * mx.t + 8), extra(8), tmp;
        pool_builder pool(pool_);
        const int S = slow_axis(dir), F = fast_axis(dir);
        const int nfast = n[F];

        for (const line& ln : m.lines(dir)) {
            if (m.dirichlet_line(ln.start.mesh_coordinate, dir, grid)) continue;
            shx_line_block b;
            std::memset(&b, 0, sizeof b);
            b.pair = ln.start.mesh_coordinate[S] * nfast + ln.start.mesh_coordinate[F];
            b.lobj = b.robj = -1;
            const int p0 = ln.start.mesh_coordinate[dir];
            const int p1 = ln.end.mesh_coordinate[dir];
            int rows = p1 - p0 + 1, columns = rows;
            int row_lo = p0, col_lo = p0;
            const int64_t base = m.ic(ln.start.mesh_coordinate) - (int64_t)p0 * m.stride(dir);

            if (const auto& o = ln.start.object) {
                const bcs::type bc = obj[o->objectID];
                const info q = st.query(bc);
                st.nbs(h, bc, o->psi, false, left.data(), extra.data());
                const int s = bc != bcs::type::Dirichlet;
                const int rL = q.r - s, tL = q.t;
                const double* lc = left.data() + s * tL;
                tmp.clear();
                for (int r = 0; r < rL; ++r)
                    for (int j = 1; j < tL; ++j) tmp.push_back(lc[r * tL + j]);
                b.nl = rL;
                b.lcols = tL - 1;
                b.lcoef = pool.add(tmp.data(), (int)tmp.size());
                tmp.clear();
                for (int r = 0; r < rL; ++r) tmp.push_back(lc[r * tL]);
                b.lbw = pool.add(tmp.data(), (int)tmp.size());
                b.lobj = (int32_t)o->object_coordinate;
                --rows;
                ++row_lo;
                --columns;
                ++col_lo;
            } else {
                const bcs::type bc = grid[dir].left;
                const info q = st.query(bc);
                st.nbs(h, bc, 1.0, false, left.data(), extra.data());
                b.nl = q.r;
                b.lcols = q.t;
                b.lcoef = pool.add(left.data(), q.r * q.t);
                if (bc == bcs::type::Dirichlet) {
                    --rows;
                    ++row_lo;
                    b.flags |= 1;
                } else if (bc == bcs::type::Neumann) {
                    for (int r = 0; r < q.nextra; ++r) {
                        neu_row_.push_back(base + (int64_t)(row_lo + r) * m.stride(dir));
                        neu_col_.push_back(base + (int64_t)row_lo * m.stride(dir));
                        neu_w_.push_back(extra[r]);
                    }
                }
            }

            if (const auto& o = ln.end.object) {
                const bcs::type bc = obj[o->objectID];
                const info q = st.query(bc);
                st.nbs(h, bc, o->psi, true, right.data(), extra.data());
                const int s = bc != bcs::type::Dirichlet;
                const int rR = q.r - s, tR = q.t;
                tmp.clear();
                for (int r = 0; r < rR; ++r)
                    for (int j = 0; j < tR - 1; ++j) tmp.push_back(right[r * tR + j]);
                b.nr = rR;
                b.rcols = tR - 1;
                b.rcoef = pool.add(tmp.data(), (int)tmp.size());
                tmp.clear();
                for (int r = 0; r < rR; ++r) tmp.push_back(right[r * tR + tR - 1]);
                b.rbw = pool.add(tmp.data(), (int)tmp.size());
                b.robj = (int32_t)o->object_coordinate;
                --rows;
                --columns;
            } else {
                const bcs::type bc = grid[dir].right;
                const info q = st.query(bc);
                st.nbs(h, bc, 1.0, true, right.data(), extra.data());
                b.nr = q.r;
                b.rcols = q.t;
                b.rcoef = pool.add(right.data(), q.r * q.t);
                if (bc == bcs::type::Dirichlet) {
                    --rows;
                    b.flags |= 2;
                } else if (bc == bcs::type::Neumann) {
                    for (int r = 0; r < q.nextra; ++r) {
                        neu_row_.push_back(base + (int64_t)(p1 + r - q.nextra + 1) * m.stride(dir));
                        neu_col_.push_back(base + (int64_t)p1 * m.stride(dir));
                        neu_w_.push_back(extra[r]);
                    }
                }
            }
            b.lo = row_lo;
            b.ni = rows - b.nl - b.nr;
            b.lcol0 = col_lo;
            b.rcol0 = col_lo + (columns - b.rcols);
            blocks_.push_back(b);
        }
    }

    shx_derivative_desc desc;
    std::memset(&desc, 0, sizeof desc);
    desc.dir = dir;
    for (int a = 0; a < 3; ++a) {
        desc.n[a] = n[a];
        desc.nR[a] = (int64_t)m.R(a).size();
    }
    desc.width = (int)interior_.size();
    for (int j = 0; j < desc.width; ++j) desc.interior[j] = interior_[j];
    desc.blocks = blocks_.data();
    desc.nblocks = (int64_t)blocks_.size();
    if (pool_.empty()) pool_.push_back(0.0);
    desc.pool = pool_.data();
    desc.npool = (int64_t)pool_.size();
    desc.neu_row = neu_row_.data();
    desc.neu_col = neu_col_.data();
    desc.neu_w = neu_w_.data();
    desc.nneu = (int64_t)neu_w_.size();
    // without a context only the host assembly is built (structure checks)
    if (ctx) status_ = shx_derivative_create(ctx, &desc, &dev_);
}

derivative::~derivative()
{
    if (dev_) shx_derivative_destroy(dev_);
}

int derivative::operator()(const shx_field& u, shx_field& du, bool plus_eq) const
{
    return shx_derivative_apply(dev_, &u, nullptr, &du, plus_eq ? 1 : 0);
}

gradient::gradient(shx_ctx* ctx, const mesh& m, const stencil& st, const bcs::Grid& grid,
                   const bcs::Object& obj)
{
    for (int i = 0; i < 3; ++i) d_[i] = std::make_unique<derivative>(ctx, i, m, st, grid, obj);
}

int gradient::status() const
{
    for (const auto& d : d_)
        if (d->status() != SHX_OK) return d->status();
    return SHX_OK;
}

int gradient::operator()(const shx_field& u, shx_field& dux, shx_field& duy, shx_field& duz) const
{
    shx_derivative* const d[3] = {d_[0]->device(), d_[1]->device(), d_[2]->device()};
    shx_field* const du[3] = {&dux, &duy, &duz};
    return shx_gradient_apply(d, &u, du);
}

laplacian::laplacian(shx_ctx* ctx, const mesh& m, const stencil& st, const bcs::Grid& grid,
                     const bcs::Object& obj)
{
    for (int i = 0; i < 3; ++i) d_[i] = std::make_unique<derivative>(ctx, i, m, st, grid, obj);
}

int laplacian::status() const
{
    for (const auto& d : d_)
        if (d->status() != SHX_OK) return d->status();
    return SHX_OK;
}

int laplacian::operator()(const shx_field& u, shx_field& du) const
{
    shx_derivative* const d[3] = {d_[0]->device(), d_[1]->device(), d_[2]->device()};
    return shx_laplacian_apply(d, &u, nullptr, &du);
}

}  // namespace ccs
