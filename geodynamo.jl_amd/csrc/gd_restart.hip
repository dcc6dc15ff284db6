// gd_restart.hip -- restart files of the step in the reference's NetCDF layout
// (outputs_writer.jl:363-409 create_netcdf_file, :410-465 setup_coordinates!,
// :466-526 setup_field_variables!, :541-631 write_*, :1221-1273
// write_restart!, :1275-1370 read_restart!), host code only.
//
// The file is NetCDF classic with 64-bit offsets (CDF-2), written and parsed
// here without a NetCDF library; every NetCDF reader opens it.  Variables
// carry the reference's names, dimensions and attributes.  Beyond the
// reference's restart, the AB2 history (the previous step's nonlinear terms,
// kept as prev_nonlinear by simulation.jl:1312-1329) is stored as
// <component>_prev_nonlinear_{real,imag}, so a run continued from the file
// takes exactly the steps the uninterrupted run takes.  A file without those
// variables (one the reference wrote) restores with an AB1 bootstrap.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "gd_internal.hpp"

namespace {

enum NcType { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
constexpr uint32_t TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C;

int type_size(int t)
{
    switch (t) {
    case NC_BYTE:
    case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT:
    case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: return 0;
    }
}

size_t pad4(size_t n) { return (n + 3) & ~size_t(3); }

// big-endian encoding into a byte vector
struct Out {
    std::vector<unsigned char> b;
    void u32(uint32_t v)
    {
        for (int k = 3; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void u64(uint64_t v)
    {
        for (int k = 7; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
    }
    void name(const std::string &s)
    {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        while (b.size() % 4) b.push_back(0);
    }
};

// an attribute: its type and the big-endian payload (unpadded)
struct Att {
    std::string name;
    int type = NC_CHAR;
    uint32_t n = 0;
    std::vector<unsigned char> bytes;
};

Att att_text(const std::string &name, const std::string &v)
{
    Att a;
    a.name = name;
    a.type = NC_CHAR;
    a.n = (uint32_t)v.size();
    a.bytes.assign(v.begin(), v.end());
    return a;
}

Att att_ints(const std::string &name, std::vector<int32_t> v)
{
    Att a;
    a.name = name;
    a.type = NC_INT;
    a.n = (uint32_t)v.size();
    Out o;
    for (int32_t x : v) o.u32((uint32_t)x);
    a.bytes = o.b;
    return a;
}

Att att_doubles(const std::string &name, std::vector<double> v)
{
    Att a;
    a.name = name;
    a.type = NC_DOUBLE;
    a.n = (uint32_t)v.size();
    Out o;
    for (double x : v) {
        uint64_t u;
        std::memcpy(&u, &x, 8);
        o.u64(u);
    }
    a.bytes = o.b;
    return a;
}

void put_atts(Out &o, const std::vector<Att> &atts)
{
    if (atts.empty()) {
        o.u32(0);
        o.u32(0);
        return;
    }
    o.u32(TAG_ATT);
    o.u32((uint32_t)atts.size());
    for (const Att &a : atts) {
        o.name(a.name);
        o.u32((uint32_t)a.type);
        o.u32(a.n);
        o.b.insert(o.b.end(), a.bytes.begin(), a.bytes.end());
        while (o.b.size() % 4) o.b.push_back(0);
    }
}

// a variable to write: native-endian host data converted on output
struct WVar {
    std::string name;
    std::vector<int> dims;
    std::vector<Att> atts;
    int type = NC_DOUBLE;
    const double *f64 = nullptr; // source for NC_DOUBLE / NC_FLOAT
    const int32_t *i32 = nullptr; // source for NC_INT
    const char *chr = nullptr;    // source for NC_CHAR
    size_t count = 0;
    uint64_t begin = 0;
};

struct Writer {
    std::vector<std::pair<std::string, uint64_t>> dims;
    std::vector<Att> gatts;
    std::vector<WVar> vars;
    int dim(const std::string &n, uint64_t len)
    {
        dims.push_back({n, len});
        return (int)dims.size() - 1;
    }

    void header(Out &o) const
    {
        o.b.clear();
        o.b.push_back('C');
        o.b.push_back('D');
        o.b.push_back('F');
        o.b.push_back(2);
        o.u32(0); // numrecs: no record dimension
        o.u32(TAG_DIM);
        o.u32((uint32_t)dims.size());
        for (const auto &d : dims) {
            o.name(d.first);
            o.u32((uint32_t)d.second);
        }
        put_atts(o, gatts);
        o.u32(TAG_VAR);
        o.u32((uint32_t)vars.size());
        for (const WVar &v : vars) {
            o.name(v.name);
            o.u32((uint32_t)v.dims.size());
            for (int d : v.dims) o.u32((uint32_t)d);
            put_atts(o, v.atts);
            o.u32((uint32_t)v.type);
            const uint64_t vs = pad4(v.count * type_size(v.type));
            o.u32(vs > 0xffffffffull ? 0xffffffffu : (uint32_t)vs);
            o.u64(v.begin);
        }
    }

    int write(const char *path)
    {
        Out o;
        header(o); // sizes only: the offsets are 8-byte fields either way
        uint64_t off = o.b.size();
        for (WVar &v : vars) {
            v.begin = off;
            off += pad4(v.count * type_size(v.type));
        }
        header(o);
        FILE *f = std::fopen(path, "wb");
        GD_CHECK(f, GD_ERR_ARG, "restart: cannot create %s", path);
        bool ok = std::fwrite(o.b.data(), 1, o.b.size(), f) == o.b.size();
        std::vector<unsigned char> buf;
        for (const WVar &v : vars) {
            const size_t es = type_size(v.type);
            const size_t chunk = 1 << 16;
            for (size_t s = 0; ok && s < v.count; s += chunk) {
                const size_t n = v.count - s < chunk ? v.count - s : chunk;
                buf.resize(n * es);
                unsigned char *p = buf.data();
                for (size_t k = 0; k < n; ++k, p += es) {
                    if (v.type == NC_DOUBLE) {
                        uint64_t u;
                        std::memcpy(&u, &v.f64[s + k], 8);
                        u = __builtin_bswap64(u);
                        std::memcpy(p, &u, 8);
                    } else if (v.type == NC_FLOAT) {
                        const float x = (float)v.f64[s + k];
                        uint32_t u;
                        std::memcpy(&u, &x, 4);
                        u = __builtin_bswap32(u);
                        std::memcpy(p, &u, 4);
                    } else if (v.type == NC_INT) {
                        const uint32_t u = __builtin_bswap32((uint32_t)v.i32[s + k]);
                        std::memcpy(p, &u, 4);
                    } else {
                        *p = (unsigned char)v.chr[s + k];
                    }
                }
                ok = ok && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
            }
            static const unsigned char zeros[4] = {0, 0, 0, 0};
            const size_t padn = pad4(v.count * es) - v.count * es;
            if (ok && padn) ok = std::fwrite(zeros, 1, padn, f) == padn;
        }
        ok = (std::fclose(f) == 0) && ok;
        GD_CHECK(ok, GD_ERR_ARG, "restart: write to %s failed", path);
        return GD_OK;
    }
};

// ---------------------------------------------------------------- reader
struct RVar {
    std::string name;
    std::vector<int> dims;
    std::map<std::string, Att> atts;
    int type = 0;
    uint64_t begin = 0;
    size_t count = 0;
};

struct Reader {
    std::vector<unsigned char> hdr; // the whole file (restart files are read in full anyway)
    size_t pos = 0;
    int version = 0;
    std::vector<std::pair<std::string, uint64_t>> dims;
    std::vector<char> recdim;
    std::map<std::string, Att> gatts;
    std::map<std::string, RVar> vars;
    std::string err;

    bool need(size_t n)
    {
        if (pos + n > hdr.size()) {
            err = "truncated header";
            return false;
        }
        return true;
    }
    bool u32(uint32_t &v)
    {
        if (!need(4)) return false;
        v = (uint32_t)hdr[pos] << 24 | (uint32_t)hdr[pos + 1] << 16 | (uint32_t)hdr[pos + 2] << 8 | hdr[pos + 3];
        pos += 4;
        return true;
    }
    bool u64(uint64_t &v)
    {
        uint32_t a, b;
        if (!u32(a) || !u32(b)) return false;
        v = (uint64_t)a << 32 | b;
        return true;
    }
    bool name(std::string &s)
    {
        uint32_t n;
        if (!u32(n) || !need(pad4(n))) return false;
        s.assign((const char *)&hdr[pos], n);
        pos += pad4(n);
        return true;
    }
    bool atts(std::map<std::string, Att> &out)
    {
        uint32_t tag, n;
        if (!u32(tag) || !u32(n)) return false;
        if (tag == 0 && n == 0) return true;
        if (tag != TAG_ATT) {
            err = "bad attribute list";
            return false;
        }
        for (uint32_t k = 0; k < n; ++k) {
            Att a;
            uint32_t t, cnt;
            if (!name(a.name) || !u32(t) || !u32(cnt)) return false;
            a.type = (int)t;
            a.n = cnt;
            const size_t bytes = (size_t)cnt * type_size(a.type);
            if (!type_size(a.type)) {
                err = "bad attribute type";
                return false;
            }
            if (!need(pad4(bytes))) return false;
            a.bytes.assign(hdr.begin() + pos, hdr.begin() + pos + bytes);
            pos += pad4(bytes);
            out[a.name] = a;
        }
        return true;
    }

    bool parse()
    {
        if (hdr.size() < 8 || hdr[0] != 'C' || hdr[1] != 'D' || hdr[2] != 'F') {
            err = "not a NetCDF classic file";
            return false;
        }
        version = hdr[3];
        if (version != 1 && version != 2) {
            err = "NetCDF format version " + std::to_string(version) + " (CDF-1 and CDF-2 are read)";
            return false;
        }
        pos = 4;
        uint32_t numrecs, tag, n;
        if (!u32(numrecs) || !u32(tag) || !u32(n)) return false;
        if (!(tag == TAG_DIM || (tag == 0 && n == 0))) {
            err = "bad dimension list";
            return false;
        }
        for (uint32_t k = 0; k < n; ++k) {
            std::string nm;
            uint32_t len;
            if (!name(nm) || !u32(len)) return false;
            dims.push_back({nm, len ? len : numrecs});
            recdim.push_back(len == 0);
        }
        if (!atts(gatts)) return false;
        if (!u32(tag) || !u32(n)) return false;
        if (!(tag == TAG_VAR || (tag == 0 && n == 0))) {
            err = "bad variable list";
            return false;
        }
        for (uint32_t k = 0; k < n; ++k) {
            RVar v;
            uint32_t nd, t, vsize;
            if (!name(v.name) || !u32(nd)) return false;
            v.count = 1;
            for (uint32_t d = 0; d < nd; ++d) {
                uint32_t id;
                if (!u32(id)) return false;
                if (id >= dims.size()) {
                    err = "bad dimension id";
                    return false;
                }
                if (recdim[id]) {
                    err = "record variable " + v.name + " (only fixed-size variables are read)";
                    return false;
                }
                v.dims.push_back((int)id);
                // element count bounded by the file size at every factor (no
                // 64-bit wrap-around from a malformed header)
                const uint64_t len = dims[id].second;
                if (len != 0 && v.count > (uint64_t)hdr.size() / len) {
                    err = "variable " + v.name + " is larger than the file";
                    return false;
                }
                v.count *= len;
            }
            if (!atts(v.atts) || !u32(t) || !u32(vsize)) return false;
            v.type = (int)t;
            if (version == 1) {
                uint32_t b;
                if (!u32(b)) return false;
                v.begin = b;
            } else if (!u64(v.begin))
                return false;
            const uint64_t ts = type_size(v.type), fsz = hdr.size();
            if (!ts || v.begin > fsz || v.count > (fsz - v.begin) / ts) {
                err = "variable " + v.name + " lies outside the file";
                return false;
            }
            vars[v.name] = v;
        }
        return true;
    }

    const RVar *var(const std::string &n) const
    {
        auto it = vars.find(n);
        return it == vars.end() ? nullptr : &it->second;
    }
    uint64_t dimlen(const std::string &n) const
    {
        for (const auto &d : dims)
            if (d.first == n) return d.second;
        return 0;
    }

    // element k of a numeric variable as double
    double at(const RVar &v, size_t k) const
    {
        const unsigned char *p = hdr.data() + v.begin + k * type_size(v.type);
        switch (v.type) {
        case NC_DOUBLE: {
            uint64_t u = 0;
            for (int j = 0; j < 8; ++j) u = u << 8 | p[j];
            double x;
            std::memcpy(&x, &u, 8);
            return x;
        }
        case NC_FLOAT: {
            uint32_t u = 0;
            for (int j = 0; j < 4; ++j) u = u << 8 | p[j];
            float x;
            std::memcpy(&x, &u, 4);
            return x;
        }
        case NC_INT: {
            uint32_t u = 0;
            for (int j = 0; j < 4; ++j) u = u << 8 | p[j];
            return (double)(int32_t)u;
        }
        case NC_SHORT: return (double)(int16_t)((uint16_t)p[0] << 8 | p[1]);
        case NC_BYTE: return (double)(signed char)p[0];
        default: return 0.0;
        }
    }
};

const char *COMPONENTS[5] = {"velocity_toroidal", "velocity_poloidal", "magnetic_toroidal", "magnetic_poloidal",
                             "temperature_spectral"};

// (l, m) pairs in the step's l-major order (shtnskit_transforms.jl:180-188)
void lm_pairs(int lmax, int mmax, std::vector<int32_t> &l, std::vector<int32_t> &m)
{
    l.clear();
    m.clear();
    for (int a = 0; a <= lmax; ++a)
        for (int b = 0; b <= (a < mmax ? a : mmax); ++b) {
            l.push_back(a);
            m.push_back(b);
        }
}

std::string geometry_of(const gd_restart_meta *meta)
{
    if (!meta || !meta->geometry[0]) return "shell";
    return std::string(meta->geometry, strnlen(meta->geometry, sizeof(meta->geometry)));
}

} // namespace

extern "C" int gd_restart_write_arrays(const char *path, const gd_restart_meta *meta, int lmax, int mmax, int nr,
                                       const double *r, int step, const double *re, const double *im,
                                       const double *nl_re, const double *nl_im, int precision,
                                       const gd_restart_comp *comp)
{
    GD_CHECK(path && re && im && r, GD_ERR_ARG, "gd_restart_write_arrays: null argument");
    GD_CHECK(lmax >= 0 && mmax >= 0 && mmax <= lmax && nr >= 2, GD_ERR_ARG,
             "gd_restart_write_arrays: lmax=%d mmax=%d nr=%d", lmax, mmax, nr);
    GD_CHECK(precision == GD_PREC_FP64 || precision == GD_PREC_FP32, GD_ERR_ARG,
             "gd_restart_write_arrays: precision must be GD_PREC_FP64 or GD_PREC_FP32");
    GD_CHECK(step >= 0, GD_ERR_ARG, "gd_restart_write_arrays: step %d < 0", step);
    GD_CHECK((nl_re == nullptr) == (nl_im == nullptr), GD_ERR_ARG,
             "gd_restart_write_arrays: give both history planes or neither");
    GD_CHECK(!nl_re || step > 0, GD_ERR_ARG, "gd_restart_write_arrays: a nonlinear history needs step > 0");
    const bool wc = comp && comp->re;
    GD_CHECK(!wc || comp->im, GD_ERR_ARG, "gd_restart_write_arrays: composition needs both planes");
    const bool wch = wc && comp->nl_re && comp->nl_im && nl_re;
    const int ft = precision == GD_PREC_FP64 ? NC_DOUBLE : NC_FLOAT;
    std::vector<int32_t> lv, mv;
    lm_pairs(lmax, mmax, lv, mv);
    const size_t nlm = lv.size(), fs = nlm * (size_t)nr;
    const std::string geom = geometry_of(meta);
    const double time = meta ? meta->time : 0.0;

    Writer w;
    // global attributes (create_netcdf_file, outputs_writer.jl:371-398)
    char stamp[64];
    const std::time_t now = std::time(nullptr);
    std::strftime(stamp, sizeof(stamp), "%Y-%m-%dT%H:%M:%S", std::gmtime(&now));
    const int rank = meta ? meta->mpi_rank : 0, nranks = meta ? meta->mpi_total_ranks : 1;
    w.gatts.push_back(att_text("title", "GeoDynamo Simulation Output"));
    w.gatts.push_back(att_text("source", "GeoDynamo simulation code"));
    w.gatts.push_back(att_text("history", std::string("Created on ") + stamp + " by rank " + std::to_string(rank)));
    w.gatts.push_back(att_text("Conventions", "CF-1.8"));
    w.gatts.push_back(att_ints("mpi_rank", {rank}));
    w.gatts.push_back(att_ints("mpi_total_ranks", {nranks}));
    // restart metadata (write_restart!, :1235-1239)
    w.gatts.push_back(att_doubles("current_time", {time}));
    w.gatts.push_back(att_ints("current_step", {step}));
    w.gatts.push_back(att_text("geometry", geom));
    w.gatts.push_back(att_doubles("restart_time", {time}));
    w.gatts.push_back(att_doubles("last_output_time", {meta ? meta->last_output_time : 0.0}));
    w.gatts.push_back(att_ints("output_count", {meta ? meta->output_count : 0}));
    w.gatts.push_back(att_ints("restart_count", {meta ? meta->restart_count : 0}));

    const int d_meta = w.dim("meta", 1);
    const int d_geo = w.dim("geometry_strlen", geom.size());
    const int d_time = w.dim("time", 1);
    const int d_r = w.dim("r", nr);
    const int d_lm = w.dim("spectral_mode", nlm);
    const int d_sc = w.dim("scalar", 1);

    const double t_arr[1] = {time};
    const int32_t step_arr[1] = {step};
    const double lot[1] = {meta ? meta->last_output_time : 0.0};
    const int32_t oc[1] = {meta ? meta->output_count : 0}, rc[1] = {meta ? meta->restart_count : 0};
    const int32_t gw[1] = {meta ? meta->grid_file_written : 0};
    auto add = [&](WVar v) { w.vars.push_back(v); };
    WVar v;
    v = WVar{};
    v.name = "geometry";
    v.dims = {d_meta, d_geo};
    v.type = NC_CHAR;
    v.chr = geom.c_str();
    v.count = geom.size();
    add(v);
    v = WVar{};
    v.name = "time";
    v.dims = {d_time};
    v.type = ft;
    v.f64 = t_arr;
    v.count = 1;
    v.atts = {att_text("long_name", "simulation_time"), att_text("units", "dimensionless")};
    add(v);
    v = WVar{};
    v.name = "step";
    v.dims = {d_time};
    v.type = NC_INT;
    v.i32 = step_arr;
    v.count = 1;
    v.atts = {att_text("long_name", "simulation_step")};
    add(v);
    v = WVar{};
    v.name = "r";
    v.dims = {d_r};
    v.type = ft;
    v.f64 = r;
    v.count = nr;
    v.atts = {att_text("long_name", "radial_coordinate"), att_text("units", "dimensionless"),
              att_doubles("valid_range", {r[0], r[nr - 1]})};
    add(v);
    v = WVar{};
    v.name = "l_values";
    v.dims = {d_lm};
    v.type = NC_INT;
    v.i32 = lv.data();
    v.count = nlm;
    v.atts = {att_text("long_name", "spherical_harmonic_degree")};
    add(v);
    v.name = "m_values";
    v.i32 = mv.data();
    v.atts = {att_text("long_name", "spherical_harmonic_order")};
    add(v);
    // spectral fields: Julia dims (spectral_mode, r) are (r, spectral_mode) in
    // the file's row-major order, the step's l-major plane [r][lm]
    auto field_vars = [&](const std::string &stem, const double *pr, const double *pi, const char *what) {
        for (int k = 0; k < 5; ++k) {
            for (int ri = 0; ri < 2; ++ri) {
                WVar f;
                f.name = std::string(COMPONENTS[k]) + stem + (ri ? "_imag" : "_real");
                f.dims = {d_r, d_lm};
                f.type = ft;
                f.f64 = (ri ? pi : pr) + (size_t)k * fs;
                f.count = fs;
                f.atts = {att_text("long_name", std::string(COMPONENTS[k]) + stem +
                                                    (ri ? "_imaginary_coefficients" : "_real_coefficients")),
                          att_text("representation", what)};
                add(f);
            }
        }
    };
    field_vars("", re, im, "spectral_space");
    if (nl_re) field_vars("_prev_nonlinear", nl_re, nl_im, "spectral_space_ab2_history");
    // composition_spectral (setup_field_variables!, outputs_writer.jl:507-522)
    auto comp_vars = [&](const std::string &stem, const double *pr, const double *pi, const char *what) {
        for (int ri = 0; ri < 2; ++ri) {
            WVar f;
            f.name = std::string("composition_spectral") + stem + (ri ? "_imag" : "_real");
            f.dims = {d_r, d_lm};
            f.type = ft;
            f.f64 = ri ? pi : pr;
            f.count = fs;
            f.atts = {att_text("long_name", std::string("composition_spectral") + stem +
                                                (ri ? "_imaginary_coefficients" : "_real_coefficients")),
                      att_text("representation", what)};
            add(f);
        }
    };
    if (wc) comp_vars("", comp->re, comp->im, "spectral_space");
    if (wch) comp_vars("_prev_nonlinear", comp->nl_re, comp->nl_im, "spectral_space_ab2_history");
    // restart scalars (write_restart!, :1252-1256)
    v = WVar{};
    v.name = "last_output_time";
    v.dims = {d_sc};
    v.type = ft;
    v.f64 = lot;
    v.count = 1;
    add(v);
    v = WVar{};
    v.name = "output_count";
    v.dims = {d_sc};
    v.type = NC_INT;
    v.i32 = oc;
    v.count = 1;
    add(v);
    v.name = "restart_count";
    v.i32 = rc;
    add(v);
    v.name = "grid_file_written";
    v.i32 = gw;
    add(v);
    return w.write(path);
}

extern "C" int gd_restart_read_arrays(const char *path, int lmax, int mmax, int nr, double *r, int *step,
                                      double *re, double *im, double *nl_re, double *nl_im, int *has_history,
                                      gd_restart_meta *meta, gd_restart_comp *comp)
{
    GD_CHECK(path && re && im && step && has_history, GD_ERR_ARG, "gd_restart_read_arrays: null argument");
    GD_CHECK(lmax >= 0 && mmax >= 0 && mmax <= lmax && nr >= 2, GD_ERR_ARG,
             "gd_restart_read_arrays: lmax=%d mmax=%d nr=%d", lmax, mmax, nr);
    Reader rd;
    {
        FILE *f = std::fopen(path, "rb");
        GD_CHECK(f, GD_ERR_ARG, "restart: cannot open %s", path);
        std::fseek(f, 0, SEEK_END);
        const long len = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        rd.hdr.resize(len > 0 ? (size_t)len : 0);
        const bool ok = len > 0 && std::fread(rd.hdr.data(), 1, rd.hdr.size(), f) == rd.hdr.size();
        std::fclose(f);
        GD_CHECK(ok, GD_ERR_ARG, "restart: cannot read %s", path);
    }
    GD_CHECK(rd.parse(), GD_ERR_ARG, "restart %s: %s", path, rd.err.c_str());
    std::vector<int32_t> lv, mv;
    lm_pairs(lmax, mmax, lv, mv);
    const size_t nlm = lv.size(), fs = nlm * (size_t)nr;
    GD_CHECK(rd.dimlen("spectral_mode") == nlm && rd.dimlen("r") == (uint64_t)nr, GD_ERR_ARG,
             "restart %s: spectral_mode=%llu r=%llu, expected %zu and %d", path,
             (unsigned long long)rd.dimlen("spectral_mode"), (unsigned long long)rd.dimlen("r"), nlm, nr);
    // file mode k -> step index: by the (l, m) variables where present
    std::vector<size_t> dest(nlm);
    for (size_t k = 0; k < nlm; ++k) dest[k] = k;
    const RVar *vl = rd.var("l_values"), *vm = rd.var("m_values");
    if (vl && vm && vl->count == nlm && vm->count == nlm) {
        std::map<std::pair<int, int>, size_t> where;
        for (size_t k = 0; k < nlm; ++k) where[{lv[k], mv[k]}] = k;
        std::vector<char> seen(nlm, 0);
        for (size_t k = 0; k < nlm; ++k) {
            const int l = (int)rd.at(*vl, k), m = (int)rd.at(*vm, k);
            auto it = where.find({l, m});
            GD_CHECK(it != where.end() && !seen[it->second], GD_ERR_ARG,
                     "restart %s: mode %zu (l=%d, m=%d) is not a distinct mode of lmax=%d mmax=%d", path, k, l, m,
                     lmax, mmax);
            seen[it->second] = 1;
            dest[k] = it->second;
        }
    }
    auto plane = [&](const std::string &name, double *out) -> int {
        const RVar *v = rd.var(name);
        GD_CHECK(v, GD_ERR_ARG, "restart %s: variable %s missing", path, name.c_str());
        GD_CHECK(v->count == fs && v->dims.size() == 2 && rd.dims[v->dims[0]].first == "r" &&
                     rd.dims[v->dims[1]].first == "spectral_mode",
                 GD_ERR_ARG, "restart %s: %s is not (r, spectral_mode)", path, name.c_str());
        GD_CHECK(v->type == NC_DOUBLE || v->type == NC_FLOAT, GD_ERR_ARG, "restart %s: %s is not floating point",
                 path, name.c_str());
        for (int i = 0; i < nr; ++i)
            for (size_t k = 0; k < nlm; ++k) out[(size_t)i * nlm + dest[k]] = rd.at(*v, (size_t)i * nlm + k);
        return GD_OK;
    };
    for (int k = 0; k < 5; ++k) {
        GD_TRY(plane(std::string(COMPONENTS[k]) + "_real", re + (size_t)k * fs));
        GD_TRY(plane(std::string(COMPONENTS[k]) + "_imag", im + (size_t)k * fs));
    }
    const RVar *vs = rd.var("step");
    *step = vs && vs->count >= 1 ? (int)rd.at(*vs, 0) : 0;
    bool hist = *step > 0;
    for (int k = 0; k < 5 && hist; ++k)
        hist = rd.var(std::string(COMPONENTS[k]) + "_prev_nonlinear_real") &&
               rd.var(std::string(COMPONENTS[k]) + "_prev_nonlinear_imag");
    *has_history = hist ? 1 : 0;
    if (hist && nl_re && nl_im)
        for (int k = 0; k < 5; ++k) {
            GD_TRY(plane(std::string(COMPONENTS[k]) + "_prev_nonlinear_real", nl_re + (size_t)k * fs));
            GD_TRY(plane(std::string(COMPONENTS[k]) + "_prev_nonlinear_imag", nl_im + (size_t)k * fs));
        }
    if (comp) {
        comp->present = rd.var("composition_spectral_real") && rd.var("composition_spectral_imag") ? 1 : 0;
        comp->has_history = comp->present && hist && rd.var("composition_spectral_prev_nonlinear_real") &&
                                    rd.var("composition_spectral_prev_nonlinear_imag")
                                ? 1
                                : 0;
        if (comp->present && comp->re && comp->im) {
            GD_TRY(plane("composition_spectral_real", comp->re));
            GD_TRY(plane("composition_spectral_imag", comp->im));
        }
        if (comp->has_history && comp->nl_re && comp->nl_im) {
            GD_TRY(plane("composition_spectral_prev_nonlinear_real", comp->nl_re));
            GD_TRY(plane("composition_spectral_prev_nonlinear_imag", comp->nl_im));
        }
    }
    if (r) {
        const RVar *v = rd.var("r");
        GD_CHECK(v && v->count == (size_t)nr, GD_ERR_ARG, "restart %s: variable r missing", path);
        for (int i = 0; i < nr; ++i) r[i] = rd.at(*v, i);
    }
    if (meta) {
        std::memset(meta, 0, sizeof(*meta));
        auto scalar = [&](const char *n, double dflt) {
            const RVar *v = rd.var(n);
            return v && v->count >= 1 ? rd.at(*v, 0) : dflt;
        };
        meta->time = scalar("time", 0.0);
        meta->last_output_time = scalar("last_output_time", 0.0);
        meta->output_count = (int)scalar("output_count", 0);
        meta->restart_count = (int)scalar("restart_count", 0);
        // files without the flag count as written (read_restart!, :1306-1312)
        meta->grid_file_written = (int)scalar("grid_file_written", 1);
        auto gi = [&](const char *n, int dflt) {
            auto it = rd.gatts.find(n);
            if (it == rd.gatts.end() || it->second.type != NC_INT || it->second.n < 1) return dflt;
            const unsigned char *p = it->second.bytes.data();
            return (int)(int32_t)((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]);
        };
        meta->mpi_rank = gi("mpi_rank", 0);
        meta->mpi_total_ranks = gi("mpi_total_ranks", 1);
        std::string g = "shell";
        const RVar *vg = rd.var("geometry");
        auto ga = rd.gatts.find("geometry");
        if (vg && vg->type == NC_CHAR) g.assign((const char *)rd.hdr.data() + vg->begin, vg->count);
        else if (ga != rd.gatts.end() && ga->second.type == NC_CHAR)
            g.assign(ga->second.bytes.begin(), ga->second.bytes.end());
        std::strncpy(meta->geometry, g.c_str(), sizeof(meta->geometry) - 1);
    }
    return GD_OK;
}

extern "C" int gd_restart_write(gd_step *s, const char *path, const gd_restart_meta *meta, int precision)
{
    GD_CHECK(s && path, GD_ERR_ARG, "gd_restart_write: null argument");
    int lmax, mmax, nr;
    GD_TRY(gd_step_shape(s, &lmax, &mmax, &nr));
    int nlm = 0;
    for (int l = 0; l <= lmax; ++l) nlm += (l < mmax ? l : mmax) + 1;
    const size_t fs = (size_t)nlm * nr, n = 5 * fs;
    std::vector<double> re(n), im(n), r(nr), nre, nim;
    GD_TRY(gd_step_get_state(s, re.data(), im.data(), GD_HOST));
    const int step = gd_step_count(s);
    if (step > 0) {
        nre.resize(n);
        nim.resize(n);
        GD_TRY(gd_step_get_history(s, nre.data(), nim.data(), GD_HOST));
    }
    GD_TRY(gd_step_radii(s, r.data()));
    // the composition, when the step carries one
    gd_restart_comp comp{};
    std::vector<double> c(4 * fs);
    const bool has_c = gd_step_has_composition(s) != 0;
    if (has_c) {
        comp.re = c.data();
        comp.im = c.data() + fs;
        GD_TRY(gd_step_get_composition(s, comp.re, comp.im, GD_HOST));
        if (step > 0) {
            comp.nl_re = c.data() + 2 * fs;
            comp.nl_im = c.data() + 3 * fs;
            GD_TRY(gd_step_get_composition_history(s, comp.nl_re, comp.nl_im, GD_HOST));
        }
    }
    return gd_restart_write_arrays(path, meta, lmax, mmax, nr, r.data(), step, re.data(), im.data(),
                                   step > 0 ? nre.data() : nullptr, step > 0 ? nim.data() : nullptr, precision,
                                   has_c ? &comp : nullptr);
}

extern "C" int gd_restart_read(gd_step *s, const char *path, gd_restart_meta *meta)
{
    GD_CHECK(s && path, GD_ERR_ARG, "gd_restart_read: null argument");
    int lmax, mmax, nr;
    GD_TRY(gd_step_shape(s, &lmax, &mmax, &nr));
    int nlm = 0;
    for (int l = 0; l <= lmax; ++l) nlm += (l < mmax ? l : mmax) + 1;
    const size_t fs = (size_t)nlm * nr, n = 5 * fs;
    std::vector<double> re(n), im(n), nre(n), nim(n), c(4 * fs);
    int step = 0, hist = 0;
    const bool has_c = gd_step_has_composition(s) != 0;
    gd_restart_comp comp{};
    comp.re = c.data();
    comp.im = c.data() + fs;
    comp.nl_re = c.data() + 2 * fs;
    comp.nl_im = c.data() + 3 * fs;
    GD_TRY(gd_restart_read_arrays(path, lmax, mmax, nr, nullptr, &step, re.data(), im.data(), nre.data(), nim.data(),
                                  &hist, meta, has_c ? &comp : nullptr));
    // the reference reads composition_spectral only where the run has one
    // (read_restart!, outputs_writer.jl:1327-1339); a step with composition
    // needs it from the file
    GD_CHECK(!has_c || comp.present, GD_ERR_ARG, "gd_restart_read: %s holds no composition_spectral", path);
    const bool h = hist && (!has_c || comp.has_history);
    // without a stored history the next step is an AB1 bootstrap, as at step 0
    GD_TRY(gd_step_restore(s, re.data(), im.data(), h ? nre.data() : nullptr, h ? nim.data() : nullptr, h ? step : 0,
                           GD_HOST));
    if (has_c)
        GD_TRY(gd_step_restore_composition(s, comp.re, comp.im, h ? comp.nl_re : nullptr, h ? comp.nl_im : nullptr,
                                           GD_HOST));
    return GD_OK;
}

extern "C" int gd_restart_filename(const char *output_dir, const char *prefix, const char *file_type,
                                   const char *geometry, double time, int rank, int time_first, char *buf,
                                   int buflen)
{
    GD_CHECK(prefix && file_type && buf && buflen > 0, GD_ERR_ARG, "gd_restart_filename: null argument");
    GD_CHECK(rank >= 0 && std::isfinite(time), GD_ERR_ARG, "gd_restart_filename: rank %d time %g", rank, time);
    // generate_filename (outputs_writer.jl:338-361): time as %.6f with '.' -> 'p'
    char ts[64];
    std::snprintf(ts, sizeof(ts), "%.6f", time);
    for (char *c = ts; *c; ++c)
        if (*c == '.') *c = 'p';
    char rk[16];
    std::snprintf(rk, sizeof(rk), "%04d", rank);
    const std::string geom = std::string("geom_") + (geometry && geometry[0] ? geometry : "shell");
    std::string name = std::string(prefix) + "_" + file_type + "_" + geom + "_";
    name += time_first ? std::string("time_") + ts + "_rank_" + rk : std::string("rank_") + rk + "_time_" + ts;
    name += ".nc";
    std::string full = output_dir && output_dir[0] ? std::string(output_dir) : std::string();
    if (!full.empty() && full.back() != '/') full += "/";
    full += name;
    GD_CHECK((int)full.size() < buflen, GD_ERR_ARG, "gd_restart_filename: buffer of %d bytes is too small", buflen);
    std::memcpy(buf, full.c_str(), full.size() + 1);
    return GD_OK;
}
