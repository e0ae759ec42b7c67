// operators.cpp -- host assembly of the derivative operators into the device
// segment table, and their application through shoccs_hip.h.
//
//   submatrix_size / domain_discretization  src/operators/derivative.cpp:259-449
//   inner_block placement of L | C | R      src/matrices/inner_block.cpp:17-40
//   derivative ctor                         src/operators/derivative.cpp:452-473
//   cut_discretization (R-space rows)       src/operators/derivative.cpp:17-257
//   gradient / laplacian                    src/operators/gradient.cpp:34-46,
//                                           src/operators/laplacian.cpp:37-58
#include "operators.hpp"

#include <algorithm>
#include <cmath>
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

struct cut_entry {
    int64_t row;
    int64_t col;
    double w;
};

// stencil::interp (src/stencils/stencil.hpp:213-270): weights of the
// interpolant at closest + y along dir within the line ln, and the two points
// (or cut points) it spans.
int interp_weights(const stencil& st, int dir, const int3& closest, double y, const line& ln,
                   double* c, int cap, boundary& lo, boundary& hi)
{
    const stencils::interp_info q = st.query_interp();
    const int ic = closest[dir];
    const int lc = ln.start.mesh_coordinate[dir], rc = ln.end.mesh_coordinate[dir];
    int before = (q.p & 1) ? (q.p - 1) / 2 : q.p / 2, after = (q.p - 1) / 2;
    if (y > 0) std::swap(before, after);
    const bool room_left = ic - lc > before || (ic - lc == before && !ln.start.object);
    const bool room_right = rc - ic > after || (rc - ic == after && !ln.end.object);
    if (room_left && room_right) {
        lo = boundary{closest, std::nullopt};
        hi = boundary{closest, std::nullopt};
        lo.mesh_coordinate[dir] -= before;
        hi.mesh_coordinate[dir] += after;
        return st.interp_interior(y, c, cap);
    }
    if (ic - lc < rc - ic) {  // the left end is nearer
        lo = ln.start;
        hi = boundary{ln.start.mesh_coordinate, std::nullopt};
        hi.mesh_coordinate[dir] += q.t - 1;
        const double psi = ln.start.object ? ln.start.object->psi : 1.0;
        if (ln.start.object && lc == ic) y += psi - 1;
        return st.interp_wall(ic - lc, y, psi, c, false, cap);
    }
    lo = boundary{ln.end.mesh_coordinate, std::nullopt};
    lo.mesh_coordinate[dir] -= q.t - 1;
    hi = ln.end;
    const double psi = ln.end.object ? ln.end.object->psi : 1.0;
    if (ln.end.object && rc == ic) y += 1 - psi;
    return st.interp_wall(rc - ic, y, psi, c, true, cap);
}

// cut_discretization (src/operators/derivative.cpp:159-257): the derivative
// along dir at the cut points R(r) of non-Dirichlet bodies.  For r == dir the
// body's own closure row; otherwise the derivative row along dir through the
// cut point's neighbourhood (interp_deriv_coefficients, :22-67), each of its
// points interpolated across direction r (OB_builder::add_interp_row,
// :87-128).  Entries go to D (on_D) or to R(r) (on_R).
void cut_rows(int r, int dir, const mesh& m, const stencil& st, const std::vector<bcs::type>& pbc,
              std::vector<cut_entry>& on_D, std::vector<cut_entry>& on_R)
{
    const auto& cuts = m.R(r);
    // the reference's quick exit tests the shapes' conditions; skipping the
    // Dirichlet points one by one gives the same rows
    if (cuts.empty() ||
        std::all_of(pbc.begin(), pbc.end(), [](bcs::type b) { return b == bcs::type::Dirichlet; }))
        return;
    const info mx = st.query_max();
    const double h = m.h(dir);
    const int64_t stride = m.stride(dir);
    const int last = m.extents()[dir] - 1;
    std::vector<double> c(mx.r * mx.t + 16, 0.0), wts(mx.t + 16, 0.0), extra(16, 0.0);
    for (int64_t row = 0; row < (int64_t)cuts.size(); ++row) {
        const mesh_object_info& o = cuts[row];
        const bcs::type bc = pbc[row];
        if (bc == bcs::type::Dirichlet) continue;
        const info q = st.query(bc);
        if (dir == r) {
            // the closure row that touches the body: the first one, or the
            // last one read backwards when the ray leaves the body
            st.nbs(h, bc, o.psi, o.ray_outside, c.data(), extra.data());
            const double* cr = o.ray_outside ? &c[(q.r - 1) * q.t] : &c[0];
            const int64_t step = o.ray_outside ? -stride : stride;
            const int64_t solid = m.ic(o.solid_coord);
            for (int i = 0; i < q.t; ++i) {
                const double w = o.ray_outside ? cr[q.t - 1 - i] : cr[i];
                if (i == 0) on_R.push_back({row, row, w});
                else on_D.push_back({row, solid + i * step, w});
            }
            continue;
        }
        const double* cl = c.data();
        int count = 0, offset = 0;
        if (std::abs(o.normal[dir]) > 1e-12) {
            // a wall closure along dir, started at the cut point
            const bool rw = o.normal[dir] < 0.0;
            st.nbs(h, bc, 1.0, rw, c.data(), extra.data());
            cl = rw ? &c[(q.r - 1) * q.t] : &c[0];
            count = q.t;
            offset = rw ? 1 - q.t : 0;
        } else {
            const int to_right = last - o.solid_coord[dir], to_left = o.solid_coord[dir];
            if (to_right >= q.p && to_left >= q.p) {
                st.interior(h, c.data());
                count = 2 * q.p + 1;
                offset = -q.p;
            } else if (to_right < to_left) {
                st.nbs(h, bc, 1.0, true, c.data(), extra.data());
                const int k = 1 + to_right;
                cl = &c[(q.r - k) * q.t];
                count = q.t;
                offset = k - q.t;
            } else {
                st.nbs(h, bc, 1.0, false, c.data(), extra.data());
                cl = &c[to_left * q.t];
                count = q.t;
                offset = -to_left;
            }
        }
        // the point nearest the cut on the fluid side (psi > 1/2: the solid one)
        int3 cp = o.solid_coord;
        if (o.psi <= 0.5) cp[r] += o.ray_outside ? -1 : 1;
        cp[dir] += offset;
        const double y = (o.ray_outside ? 1.0 : -1.0) * (o.psi <= 0.5 ? o.psi : o.psi - 1);
        for (int k = 0; k < count; ++k, ++cp[dir]) {
            const double v = cl[k];
            if (cp[dir] == o.solid_coord[dir]) {
                on_R.push_back({row, row, v});
                continue;
            }
            const line ln = m.interp_line(r, cp);
            boundary lo, hi;
            interp_weights(st, r, cp, y, ln, wts.data(), mx.t, lo, hi);
            const int64_t lic = m.ic(lo.mesh_coordinate), ric = m.ic(hi.mesh_coordinate);
            int e = 0;
            if (lo.object) on_R.push_back({row, lo.object->object_coordinate, v * wts[e]});
            else on_D.push_back({row, lic, v * wts[e]});
            ++e;
            for (int64_t x = lic + ln.stride; x < ric; x += ln.stride) on_D.push_back({row, x, v * wts[e++]});
            if (hi.object) on_R.push_back({row, hi.object->object_coordinate, v * wts[e]});
            else on_D.push_back({row, ric, v * wts[e]});
        }
    }
}

}  // namespace

namespace {
bcs::Points shape_conditions(const mesh& m, const bcs::Object& obj)
{
    bcs::Points p;
    point_conditions(m, obj, nullptr, p);
    return p;
}
}  // namespace

derivative::derivative(shx_ctx* ctx, int dir, const mesh& m, const stencil& st,
                       const bcs::Grid& grid, const bcs::Object& obj)
    : derivative(ctx, dir, m, st, grid, shape_conditions(m, obj))
{
}

derivative::derivative(shx_ctx* ctx, int dir, const mesh& m, const stencil& st,
                       const bcs::Grid& grid, const bcs::Points& pbc)
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
                const bcs::type bc = pbc[dir][o->object_coordinate];
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
                const bcs::type bc = pbc[dir][o->object_coordinate];
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
            if (b.ni < 0 && status_ == SHX_OK) {
                // the closures of a segment shorter than its two closure
                // blocks would overlap: the reference builds a circulant with
                // a negative row count there (derivative.cpp:428-436), which
                // is undefined, so such meshes are refused
                status_ = SHX_EUNSUPPORTED;
                error_ = "derivative along " + std::string(1, "xyz"[dir]) + ": a line segment of " +
                         std::to_string(rows) + " rows (positions " + std::to_string(p0) + ".." +
                         std::to_string(p1) + ") is shorter than its closures (" + std::to_string(b.nl) +
                         " + " + std::to_string(b.nr) + " rows); refine the mesh or move the body";
            }
            b.lcol0 = col_lo;
            b.rcol0 = col_lo + (columns - b.rcols);
            blocks_.push_back(b);
        }
    }

    // R-space rows, one row list over R(x), R(y), R(z): within a row the D
    // entries, then the R entries, each ordered by (column, weight).  Columns
    // are clamped to their buffer: a scheme without interpolants gives a
    // degenerate line whose zero-weight ends may lie past the mesh.
    rptr_.assign(1, 0);
    if (n[dir] >= 2) {
        for (int r = 0; r < 3; ++r) {
            std::vector<cut_entry> on_D, on_R;
            cut_rows(r, dir, m, st, pbc[r], on_D, on_R);
            auto order = [](const cut_entry& a, const cut_entry& b) {
                if (a.row != b.row) return a.row < b.row;
                if (a.col != b.col) return a.col < b.col;
                return a.w < b.w;
            };
            std::sort(on_D.begin(), on_D.end(), order);
            std::sort(on_R.begin(), on_R.end(), order);
            const int64_t nD = m.size(), nRr = (int64_t)m.R(r).size();
            size_t a = 0, b = 0;
            while (a < on_D.size() || b < on_R.size()) {
                const int64_t row = std::min(a < on_D.size() ? on_D[a].row : INT64_MAX,
                                             b < on_R.size() ? on_R[b].row : INT64_MAX);
                for (; a < on_D.size() && on_D[a].row == row; ++a) {
                    rcol_.push_back(std::clamp<int64_t>(on_D[a].col, 0, nD - 1));
                    rw_.push_back(on_D[a].w);
                }
                for (; b < on_R.size() && on_R[b].row == row; ++b) {
                    rcol_.push_back(-1 - std::clamp<int64_t>(on_R[b].col, 0, nRr - 1));
                    rw_.push_back(on_R[b].w);
                }
                rset_.push_back(r);
                rrow_.push_back(row);
                rptr_.push_back((int64_t)rw_.size());
            }
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
    desc.nrrow = (int64_t)rrow_.size();
    desc.rset = rset_.data();
    desc.rrow = rrow_.data();
    desc.rptr = rptr_.data();
    desc.rcol = rcol_.data();
    desc.rw = rw_.data();
    // without a context only the host assembly is built (structure checks)
    if (ctx && status_ == SHX_OK) status_ = shx_derivative_create(ctx, &desc, &dev_);
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
