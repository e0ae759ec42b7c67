// field_io.cpp -- the reference's field output: raw FP64 dumps of each scalar
// (D, then Rx, Ry, Rz), the cut-point geometry files and the XDMF3 time
// series that describes them, written when the dump interval fires.
//
//   interval / d_interval                 src/io/interval.hpp:10-103
//   field_io::write                        src/io/field_io.cpp:32-65
//   field_data::write / write_geom         src/io/field_data.cpp:8-57
//   xdmf::write (header, appended grids)   src/io/xdmf.cpp:19-165
//   write_scalar_error (the Error field)   src/systems/detail/scalar_system_utils.hpp:224-301
//
// The XDMF document is kept in memory and written whole at every dump, laid
// out as pugixml saves the reference's (tab indentation, one element per
// line, the declaration and doctype first).
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <new>
#include <optional>
#include <string>
#include <vector>

#include "operators.hpp"
#include "shoccs.h"

namespace ccs {

namespace {

namespace fs = std::filesystem;

// interval<T> (interval.hpp:10-55): fires when val - first >= distance - tol
template <typename T>
struct interval {
    std::optional<T> distance;
    T first{};
    bool operator()(T val, T tol = T{}) const { return distance && val - first >= *distance - tol; }
    explicit operator bool() const { return distance.has_value(); }
    void advance() { first += distance.value_or(T{}); }
};

// shortest round-trip form of a double, as fmt's "{}" prints it
std::string shortest(double v)
{
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, v);
    return std::string(buf, r.ptr);
}

// pugixml's attribute form of a double ("%.17g")
std::string g17(double v)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

struct grid_entry {
    int number;
    double time;
    std::vector<std::string> vars, files;
};

}  // namespace

struct field_io {
    // d_interval (interval.hpp:57-103)
    interval<int> step_iv;
    interval<double> time_iv;
    int ndumps = 0;
    bool step_ready = false, time_ready = false;

    fs::path dir;
    std::string xmf;
    int suffix = 6;
    int3 n{1, 1, 1};
    real3 lo{}, hi{};
    std::array<std::vector<real3>, 3> rpos;  // cut-point positions
    std::vector<grid_entry> grids;

    bool fires(int step, double time, double dt)
    {
        if (!(step_iv || time_iv)) return false;
        step_ready = step_iv(step);
        time_ready = time_iv(time, dt);
        return step_ready || time_ready || step == 0;
    }
    void advance()
    {
        if (step_ready) step_iv.advance();
        if (time_ready) time_iv.advance();
        ++ndumps;
    }

    std::string header() const
    {
        auto join3 = [](const real3& v) { return shortest(v[0]) + " " + shortest(v[1]) + " " + shortest(v[2]); };
        real3 d;
        for (int a = 0; a < 3; ++a) d[a] = (hi[a] - lo[a]) / std::max(n[a] - 1.0, 1.0);
        const std::string dims = std::to_string(n[0]) + " " + std::to_string(n[1]) + " " + std::to_string(n[2]);
        std::string s;
        s += "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n";
        s += "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n";
        s += "<Xdmf Version=\"3.0\">\n";
        s += "\t<Domain>\n";
        s += "\t\t<Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"" + dims + "\" />\n";
        s += "\t\t<Geometry GeometryType=\"Origin_DxDyDz\">\n";
        s += "\t\t\t<DataItem Format=\"XML\" NumberType=\"Float\" Dimensions=\"3\">" + join3(lo) + "</DataItem>\n";
        s += "\t\t\t<DataItem Format=\"XML\" NumberType=\"Float\" Dimensions=\"3\">" + join3(d) + "</DataItem>\n";
        s += "\t\t</Geometry>\n";
        const char* rname[3] = {"rx", "ry", "rz"};
        for (int a = 0; a < 3; ++a) {
            const std::string cnt = std::to_string(rpos[a].size());
            s += "\t\t<Topology TopologyType=\"Polyvertex\" NumberOfElements=\"" + cnt + "\" />\n";
            s += "\t\t<Geometry GeometryType=\"XYZ\">\n";
            s += "\t\t\t<DataItem NumberType=\"Float\" Precision=\"8\" Format=\"Binary\" Dimensions=\"" + cnt +
                 " 3\">" + rname[a] + "</DataItem>\n";
            s += "\t\t</Geometry>\n";
        }
        return s;
    }

    // append_xdmf (xdmf.cpp:19-84): one spatial collection per dump, a
    // uniform grid for D ("F") and one per non-empty cut-point set, each
    // naming the variables' files; the cut-point values follow D in the same
    // file, so their DataItems seek past what precedes them
    std::string grid(const grid_entry& g) const
    {
        std::string s;
        s += "\t\t\t<Grid Name=\"" + std::to_string(g.number) +
             "\" GridType=\"Collection\" CollectionType=\"Spatial\">\n";
        s += "\t\t\t\t<Time Value=\"" + g17(g.time) + "\" />\n";
        unsigned long offset = 0;
        auto sub = [&](unsigned long count, const char* name, int topo, const std::string& dims) {
            if (count == 0) return;
            s += std::string("\t\t\t\t<Grid Name=\"") + name + "\" GridType=\"Uniform\">\n";
            s += "\t\t\t\t\t<Topology Reference=\"/Xdmf/Domain/Topology[" + std::to_string(topo) + "]\" />\n";
            s += "\t\t\t\t\t<Geometry Reference=\"/Xdmf/Domain/Geometry[" + std::to_string(topo) + "]\" />\n";
            for (size_t i = 0; i < g.vars.size(); ++i) {
                s += "\t\t\t\t\t<Attribute Name=\"" + g.vars[i] + "\">\n";
                s += "\t\t\t\t\t\t<DataItem Dimensions=\"" + dims +
                     "\" NumberType=\"Float\" Precision=\"8\" Format=\"Binary\"";
                if (offset > 0) s += " Seek=\"" + std::to_string(offset) + "\"";
                s += ">" + g.files[i] + "</DataItem>\n";
                s += "\t\t\t\t\t</Attribute>\n";
            }
            s += "\t\t\t\t</Grid>\n";
            offset += count * sizeof(double);
        };
        const unsigned long N = (unsigned long)n[0] * n[1] * n[2];
        sub(N, "F", 1, std::to_string(n[0]) + " " + std::to_string(n[1]) + " " + std::to_string(n[2]));
        const char* gname[3] = {"RX", "RY", "RZ"};
        for (int a = 0; a < 3; ++a)
            sub(rpos[a].size(), gname[a], 2 + a, std::to_string(rpos[a].size()) + " 1");
        s += "\t\t\t</Grid>\n";
        return s;
    }

    bool save_xdmf() const
    {
        std::string s = header();
        s += "\t\t<Grid Name=\"TimeSeries\" GridType=\"Collection\" CollectionType=\"Temporal\"";
        if (grids.empty()) {
            s += " />\n";
        } else {
            s += ">\n";
            for (const auto& g : grids) s += grid(g);
            s += "\t\t</Grid>\n";
        }
        s += "\t</Domain>\n</Xdmf>\n";
        std::ofstream o(dir / xmf, std::ios::binary);
        o << s;
        return (bool)o;
    }
};

}  // namespace ccs

struct shoccs_io : ccs::field_io {
};

using namespace ccs;

extern "C" {

int shoccs_io_create(const shoccs_problem* p, const shoccs_io_desc* d, shoccs_io** out)
{
    if (!p || !d || !out) return SHX_EINVAL;
    *out = nullptr;
    auto io = std::unique_ptr<shoccs_io>(new (std::nothrow) shoccs_io());
    if (!io) return SHX_ENOMEM;
    if (d->write_every_step > 0) io->step_iv.distance = d->write_every_step;
    if (d->write_every_time > 0.0) io->time_iv.distance = d->write_every_time;
    io->dir = d->dir && *d->dir ? d->dir : "io";
    io->xmf = d->xdmf_filename && *d->xdmf_filename ? d->xdmf_filename : "view.xmf";
    io->suffix = d->suffix_length > 0 ? d->suffix_length : 6;
    int32_t n[3];
    double lo[3], hi[3];
    if (shoccs_problem_geometry(p, n, lo, hi) != SHX_OK) return SHX_EINVAL;
    for (int a = 0; a < 3; ++a) {
        io->n[a] = n[a];
        io->lo[a] = lo[a];
        io->hi[a] = hi[a];
        const int64_t m = shoccs_problem_size(p, a + 1);
        std::vector<double> pos((size_t)std::max<int64_t>(m, 1) * 3);
        if (m > 0 && shoccs_mesh_r(p, a, nullptr, pos.data(), nullptr, nullptr, nullptr) != SHX_OK)
            return SHX_EINVAL;
        for (int64_t i = 0; i < m; ++i) io->rpos[a].push_back({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
    }
    *out = io.release();
    return SHX_OK;
}

int shoccs_io_destroy(shoccs_io* io)
{
    if (!io) return SHX_EINVAL;
    delete io;
    return SHX_OK;
}

int shoccs_io_write(shoccs_io* io, int32_t nvars, const char* const* names, const double* const* fields,
                    int32_t step, double time, double dt, int32_t* wrote)
{
    if (!io || nvars < 0 || (nvars > 0 && (!names || !fields))) return SHX_EINVAL;
    if (wrote) *wrote = 0;
    if (!io->fires(step, time, dt)) return SHX_OK;
    // the reference creates the directory at step 0 only (its later writes
    // fail silently without it); here every dump makes sure it exists
    std::error_code ec;
    fs::create_directories(io->dir, ec);
    if (ec) return SHX_EINVAL;
    const int k = io->ndumps;
    grid_entry g{k, time, {}, {}};
    for (int v = 0; v < nvars; ++v) {
        char suffix[32];
        std::snprintf(suffix, sizeof suffix, "%0*d", io->suffix, k);
        g.vars.push_back(names[v]);
        g.files.push_back(std::string(names[v]) + "." + suffix);
    }
    io->grids.push_back(g);
    if (!io->save_xdmf()) return SHX_EINVAL;
    // field_data::write_geom at the first dump: x, y, z per cut point (z, y,
    // x on a mesh with one z point)
    if (k == 0) {
        const char* gname[3] = {"rx", "ry", "rz"};
        for (int a = 0; a < 3; ++a) {
            std::ofstream o(io->dir / gname[a], std::ios::binary);
            for (const real3& r : io->rpos[a]) {
                const double w[3] = {io->n[2] == 1 ? r[2] : r[0], r[1], io->n[2] == 1 ? r[0] : r[2]};
                o.write(reinterpret_cast<const char*>(w), sizeof w);
            }
            if (!o) return SHX_EINVAL;
        }
    }
    // field_data::write: D, then Rx, Ry, Rz, native FP64
    const size_t N = (size_t)io->n[0] * io->n[1] * io->n[2];
    const size_t total = N + io->rpos[0].size() + io->rpos[1].size() + io->rpos[2].size();
    for (int v = 0; v < nvars; ++v) {
        std::ofstream o(io->dir / g.files[v], std::ios::binary);
        o.write(reinterpret_cast<const char*>(fields[v]), (std::streamsize)(total * sizeof(double)));
        if (!o) return SHX_EINVAL;
    }
    io->advance();
    if (wrote) *wrote = 1;
    return SHX_OK;
}

}  // extern "C"
