// NetCDF classic-format files around the path (host code).
//
// Concentration output.  Wacomm::save (Wacomm.cpp:913-1020) writes, through
// the NetCDF library, the window's time axis, s_rho, lat/lon/mask/h on the rho
// grid, zeta and the concentration frames conc[ocean_time][s_rho][eta_rho]
// [eta_xi] with their attributes and a _FillValue of 9.99999993e36.  This
// writer produces the same dimensions, variables and attributes in the NetCDF
// classic format with 64-bit offsets (CDF-2).
//
// Particle restart.  Particles::saveAsNetCDF / defineNetCDF / putNetCDF
// (Particles.cpp:150-264) write restart version 3: provenance and restart
// attributes, dimensions particles and particle_time, id (uint64),
// object_type (uint), drift_side (int) and lat, lon, depth, k, j, i, health,
// age, time [particles][particle_time] (double).  The reference creates a
// NetCDF-4 (HDF5) file; HDF5 cannot be produced without its library, so this
// writer uses the classic format with 64-bit data (CDF-5), the classic variant
// that holds uint64, which the NetCDF library opens with the same API calls
// (Particles::loadFromNetCDF reads it unchanged).  The reader follows
// Particles::loadFromNetCDF (Particles.cpp:373-480) for CDF-1, CDF-2 and
// CDF-5 files.
//
// Format (NetCDF Classic and 64-bit Offset Format Specification; CDF-5 as
// defined by PnetCDF): magic "CDF" + version byte; big-endian header of
// dimension, attribute and variable lists; each variable's data at its begin
// offset, padded to 4 bytes.  In CDF-5 every count and length (NON_NEG) is 8
// bytes; in CDF-1 and CDF-2 it is 4; begin is 8 bytes in CDF-2 and CDF-5.
// A dimension of length 0 is the record (unlimited) dimension.
//
// Files are written as <path>.partial and renamed when complete, so a reader
// never sees a partial file (the reference's output-publication rule).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wacomm_hip.h"

namespace {

enum {
    NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6,
    NC_UBYTE = 7, NC_USHORT = 8, NC_UINT = 9, NC_INT64 = 10, NC_UINT64 = 11
};
constexpr uint32_t kDimTag = 0x0A, kVarTag = 0x0B, kAttTag = 0x0C;
constexpr const char *kRandomKeyScheme = "ordered-splitmix64-v2";     // NumericalHelpers.hpp:93

int type_size(int t) {
    switch (t) {
    case NC_BYTE: case NC_CHAR: case NC_UBYTE: return 1;
    case NC_SHORT: case NC_USHORT: return 2;
    case NC_INT: case NC_FLOAT: case NC_UINT: return 4;
    case NC_DOUBLE: case NC_INT64: case NC_UINT64: return 8;
    default: return 0;
    }
}

struct Buf {
    bool v5 = false;            // CDF-5: 8-byte counts
    std::vector<uint8_t> b;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void nonneg(uint64_t v) {
        if (v5) u64(v); else u32((uint32_t)v);
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        nonneg(s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
};

struct Att {
    std::string name;
    int type;
    std::vector<uint8_t> payload;   // big-endian values
    uint64_t n;
};

Att text(const std::string &name, const std::string &v) {
    return Att{name, NC_CHAR, std::vector<uint8_t>(v.begin(), v.end()), (uint64_t)v.size()};
}
Att f32(const std::string &name, float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    Buf b;
    b.u32(u);
    return Att{name, NC_FLOAT, b.b, 1};
}
Att f64(const std::string &name, double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    Buf b;
    b.u64(u);
    return Att{name, NC_DOUBLE, b.b, 1};
}
Att u64(const std::string &name, uint64_t v) {
    Buf b;
    b.u64(v);
    return Att{name, NC_UINT64, b.b, 1};
}

struct Var {
    std::string name;
    std::vector<int> dims;
    int type;                   // NC_FLOAT, NC_DOUBLE, NC_INT, NC_UINT or NC_UINT64
    std::vector<Att> atts;
    const void *data;           // host values in native order, or nullptr (written as `fill`)
    double fill;
    uint64_t count;
    uint64_t begin = 0;
};

void put_atts(Buf &h, const std::vector<Att> &atts) {
    if (atts.empty()) {
        h.u32(0);
        h.nonneg(0);
        return;
    }
    h.u32(kAttTag);
    h.nonneg(atts.size());
    for (const Att &a : atts) {
        h.name(a.name);
        h.u32((uint32_t)a.type);
        h.nonneg(a.n);
        h.b.insert(h.b.end(), a.payload.begin(), a.payload.end());
        h.pad();
    }
}

uint64_t var_bytes(const Var &v) { return v.count * type_size(v.type); }

// The header (numrecs 0).  A variable whose first dimension has length 0 is a
// record variable and holds no records.
Buf header(int version, const std::vector<std::pair<std::string, uint64_t>> &dims, const std::vector<Att> &gatts,
           const std::vector<Var> &vars) {
    Buf h;
    h.v5 = version == 5;
    h.b = {'C', 'D', 'F', (uint8_t)version};
    h.nonneg(0);                                // numrecs
    h.u32(kDimTag);
    h.nonneg(dims.size());
    for (const auto &d : dims) {
        h.name(d.first);
        h.nonneg(d.second);
    }
    put_atts(h, gatts);
    h.u32(kVarTag);
    h.nonneg(vars.size());
    for (const Var &v : vars) {
        h.name(v.name);
        h.nonneg(v.dims.size());
        for (int d : v.dims) h.nonneg((uint64_t)d);
        put_atts(h, v.atts);
        h.u32((uint32_t)v.type);
        // vsize: the variable's bytes, or one record's bytes for a record
        // variable (the product of its other dimensions)
        uint64_t per = v.count;
        if (!v.dims.empty() && dims[v.dims[0]].second == 0) {
            per = 1;
            for (size_t q = 1; q < v.dims.size(); ++q) per *= dims[v.dims[q]].second;
        }
        const uint64_t vsize = (per * type_size(v.type) + 3) / 4 * 4;
        if (h.v5)
            h.u64(vsize);
        else
            h.u32(vsize > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)vsize);
        h.u64(v.begin);
    }
    return h;
}

bool write_data(FILE *fp, const Var &v) {
    const size_t chunk = 1 << 18;
    const int sz = type_size(v.type);
    std::vector<uint8_t> out;
    for (uint64_t q0 = 0; q0 < v.count; q0 += chunk) {
        const uint64_t m = v.count - q0 < chunk ? v.count - q0 : chunk;
        out.resize(m * sz);
        uint8_t *o = out.data();
        for (uint64_t q = 0; q < m; ++q) {
            uint64_t u = 0;
            switch (v.type) {
            case NC_DOUBLE: {
                const double x = v.data ? ((const double *)v.data)[q0 + q] : v.fill;
                std::memcpy(&u, &x, 8);
                break;
            }
            case NC_FLOAT: {
                const float x = v.data ? ((const float *)v.data)[q0 + q] : (float)v.fill;
                uint32_t w;
                std::memcpy(&w, &x, 4);
                u = w;
                break;
            }
            case NC_UINT64: u = v.data ? ((const uint64_t *)v.data)[q0 + q] : 0; break;
            case NC_UINT: u = v.data ? ((const uint32_t *)v.data)[q0 + q] : 0; break;
            case NC_INT: u = v.data ? (uint32_t)((const int32_t *)v.data)[q0 + q] : 0; break;
            }
            for (int s = 8 * (sz - 1); s >= 0; s -= 8) *o++ = (uint8_t)(u >> s);
        }
        if (fwrite(out.data(), 1, out.size(), fp) != out.size()) return false;
    }
    const uint64_t bytes = var_bytes(v);
    static const uint8_t zero[4] = {0, 0, 0, 0};
    return bytes % 4 == 0 || fwrite(zero, 1, 4 - bytes % 4, fp) == 4 - bytes % 4;
}

// Lays out the variables after the header, writes <path>.partial, renames.
int write_file(const char *path, int version, const std::vector<std::pair<std::string, uint64_t>> &dims,
               const std::vector<Att> &gatts, std::vector<Var> &vars) {
    // two passes: the header's size fixes each variable's begin offset
    uint64_t off = header(version, dims, gatts, vars).b.size();
    for (Var &v : vars) {
        v.begin = off;
        off += (var_bytes(v) + 3) / 4 * 4;
    }
    const Buf h = header(version, dims, gatts, vars);
    const std::string tmp = std::string(path) + ".partial";
    FILE *fp = fopen(tmp.c_str(), "wb");
    if (!fp) return WC_ERR_IO;
    bool ok = fwrite(h.b.data(), 1, h.b.size(), fp) == h.b.size();
    for (const Var &v : vars) ok = ok && write_data(fp, v);
    ok = (fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return WC_ERR_IO;
    }
    return WC_OK;
}

// ---- reader ----------------------------------------------------------------

struct RAtt {
    std::string name;
    int type = 0;
    uint64_t n = 0;
    std::vector<uint8_t> values;    // big-endian
};
struct RVar {
    std::string name;
    std::vector<uint64_t> dims;
    int type = 0;
    uint64_t vsize = 0, begin = 0;
    bool record = false;
};
struct NcFile {
    std::vector<uint8_t> bytes;
    int version = 0;
    uint64_t numrecs = 0, recsize = 0;
    std::vector<std::pair<std::string, uint64_t>> dims;
    std::vector<RAtt> gatts;
    std::vector<RVar> vars;

    const RAtt *att(const char *name) const {
        for (const RAtt &a : gatts)
            if (a.name == name) return &a;
        return nullptr;
    }
    const RVar *var(const char *name) const {
        for (const RVar &v : vars)
            if (v.name == name) return &v;
        return nullptr;
    }
};

struct Reader {
    const std::vector<uint8_t> &b;
    size_t at = 0;
    bool v5 = false, ok = true;
    bool need(uint64_t n) {
        if (!ok || n > b.size() - at) ok = false;
        return ok;
    }
    uint32_t u32() {
        if (!need(4)) return 0;
        uint32_t v = 0;
        for (int q = 0; q < 4; ++q) v = v << 8 | b[at++];
        return v;
    }
    uint64_t u64() {
        if (!need(8)) return 0;
        uint64_t v = 0;
        for (int q = 0; q < 8; ++q) v = v << 8 | b[at++];
        return v;
    }
    uint64_t nonneg() { return v5 ? u64() : u32(); }
    void skip_pad() {
        while (ok && at % 4) {
            if (!need(1)) return;
            ++at;
        }
    }
    std::string name() {
        const uint64_t n = nonneg();
        if (!need(n)) return {};
        std::string s(b.begin() + at, b.begin() + at + n);
        at += n;
        skip_pad();
        return s;
    }
    std::vector<RAtt> atts() {
        std::vector<RAtt> out;
        const uint32_t tag = u32();
        const uint64_t n = nonneg();
        if (tag == 0 && n == 0) return out;
        if (tag != kAttTag) {
            ok = false;
            return out;
        }
        for (uint64_t q = 0; q < n && ok; ++q) {
            RAtt a;
            a.name = name();
            a.type = (int)u32();
            a.n = nonneg();
            const int sz = type_size(a.type);
            if (!sz || a.n > b.size() || !need(a.n * sz)) {
                ok = false;
                break;
            }
            a.values.assign(b.begin() + at, b.begin() + at + a.n * sz);
            at += a.n * sz;
            skip_pad();
            out.push_back(std::move(a));
        }
        return out;
    }
};

// Parses the header; 0 or an error code with the reason in `why`.
int parse(NcFile &f, std::string &why) {
    const auto &b = f.bytes;
    if (b.size() >= 4 && b[0] == 0x89 && b[1] == 'H' && b[2] == 'D' && b[3] == 'F') {
        why = "NetCDF-4 (HDF5) files need the NetCDF library; classic (CDF-1, CDF-2, CDF-5) files are read";
        return WC_ERR_UNSUPPORTED;
    }
    if (b.size() < 4 || b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2 && b[3] != 5)) {
        why = "not a NetCDF classic file";
        return WC_ERR_IO;
    }
    f.version = b[3];
    Reader r{b, 4};
    r.v5 = f.version == 5;
    f.numrecs = r.nonneg();
    const uint32_t dtag = r.u32();
    const uint64_t nd = r.nonneg();
    if (!(dtag == 0 && nd == 0) && dtag != kDimTag) r.ok = false;
    for (uint64_t q = 0; q < nd && r.ok && (dtag == kDimTag); ++q) {
        std::string n = r.name();
        const uint64_t len = r.nonneg();
        f.dims.emplace_back(std::move(n), len);
    }
    f.gatts = r.atts();
    const uint32_t vtag = r.u32();
    const uint64_t nv = r.nonneg();
    if (!(vtag == 0 && nv == 0) && vtag != kVarTag) r.ok = false;
    for (uint64_t q = 0; q < nv && r.ok && vtag == kVarTag; ++q) {
        RVar v;
        v.name = r.name();
        const uint64_t ndims = r.nonneg();
        for (uint64_t d = 0; d < ndims && r.ok; ++d) {
            const uint64_t id = r.nonneg();
            if (id >= f.dims.size()) r.ok = false;
            v.dims.push_back(id);
        }
        r.atts();
        v.type = (int)r.u32();
        v.vsize = r.nonneg();
        v.begin = f.version == 1 ? r.u32() : r.u64();
        if (!type_size(v.type)) r.ok = false;
        v.record = r.ok && !v.dims.empty() && f.dims[v.dims[0]].second == 0;
        f.vars.push_back(std::move(v));
    }
    if (!r.ok) {
        why = "truncated or malformed NetCDF header";
        return WC_ERR_IO;
    }
    // record size: the sum of the record variables' vsize; a single record
    // variable is not padded
    int nrec = 0;
    for (const RVar &v : f.vars)
        if (v.record) {
            ++nrec;
            f.recsize += v.vsize;
        }
    if (nrec == 1)
        for (const RVar &v : f.vars)
            if (v.record) {
                uint64_t per = type_size(v.type);
                for (size_t d = 1; d < v.dims.size(); ++d) per *= f.dims[v.dims[d]].second;
                f.recsize = per;
            }
    return WC_OK;
}

std::string att_text(const RAtt &a) {
    if (a.type != NC_CHAR) return std::string();
    std::string s(a.values.begin(), a.values.end());
    while (!s.empty() && s.back() == '\0') s.pop_back();
    return s;
}

// The element at `q` of a numeric attribute or variable, as double and as
// integer (exact for 64-bit integers).
struct Num {
    double d;
    uint64_t u;
    bool neg, integral;
};
Num read_num(const uint8_t *p, int type) {
    uint64_t raw = 0;
    const int sz = type_size(type);
    for (int q = 0; q < sz; ++q) raw = raw << 8 | p[q];
    Num n{0, 0, false, true};
    switch (type) {
    case NC_BYTE: { const int8_t v = (int8_t)raw; n.d = v; n.neg = v < 0; n.u = (uint64_t)(int64_t)v; break; }
    case NC_UBYTE: case NC_USHORT: case NC_UINT: case NC_UINT64: n.u = raw; n.d = (double)raw; break;
    case NC_SHORT: { const int16_t v = (int16_t)raw; n.d = v; n.neg = v < 0; n.u = (uint64_t)(int64_t)v; break; }
    case NC_INT: { const int32_t v = (int32_t)raw; n.d = v; n.neg = v < 0; n.u = (uint64_t)(int64_t)v; break; }
    case NC_INT64: { const int64_t v = (int64_t)raw; n.d = (double)v; n.neg = v < 0; n.u = (uint64_t)v; break; }
    case NC_FLOAT: {
        float v;
        const uint32_t w = (uint32_t)raw;
        std::memcpy(&v, &w, 4);
        n.d = v;
        break;
    }
    case NC_DOUBLE: std::memcpy(&n.d, &raw, 8); break;
    }
    if (type == NC_FLOAT || type == NC_DOUBLE) {
        n.neg = n.d < 0;
        n.integral = std::isfinite(n.d) && n.d == std::floor(n.d) && std::fabs(n.d) < 18446744073709551616.0;
        n.u = n.integral ? (n.neg ? (uint64_t)(int64_t)n.d : (uint64_t)n.d) : 0;
    }
    return n;
}

// Element offset of variable v at flat index q of its first dimension with
// the other indices 0 (getVar({q,0,..},{1,1,..})).  False when the element
// does not lie inside the file; every product and sum of header values is
// checked for overflow, so a corrupt header cannot wrap an offset back into
// range.
bool element_offset(const NcFile &f, const RVar &v, uint64_t q, uint64_t &off) {
    const uint64_t sz = (uint64_t)type_size(v.type);
    uint64_t inner = 1, stride = 0, rel = 0;
    for (size_t d = 1; d < v.dims.size(); ++d)
        if (__builtin_mul_overflow(inner, f.dims[v.dims[d]].second, &inner)) return false;
    if (v.record) stride = f.recsize;
    else if (__builtin_mul_overflow(inner, sz, &stride)) return false;
    if (__builtin_mul_overflow(q, stride, &rel) || __builtin_add_overflow(v.begin, rel, &off)) return false;
    return sz > 0 && off <= f.bytes.size() && sz <= f.bytes.size() - off;
}

uint64_t first_dim_size(const NcFile &f, const RVar &v) {
    if (v.dims.empty()) return 1;
    return v.record ? f.numrecs : f.dims[v.dims[0]].second;
}

void set_err(char *err, size_t len, const std::string &msg) {
    if (!err || !len) return;
    const size_t n = msg.size() < len - 1 ? msg.size() : len - 1;
    std::memcpy(err, msg.data(), n);
    err[n] = '\0';
}

}  // namespace

extern "C" int wc_write_concentration_nc(const char *path, const wc_conc_file *f) {
    if (!path || !f || f->T < 1 || f->s_rho < 1 || f->eta < 1 || f->xi < 1 || !f->ocean_time || !f->s_rho_values ||
        !f->lat_rho || !f->lon_rho || !f->mask_rho || !f->h)
        return WC_ERR_INVALID;
    const uint64_t T = f->T, S = f->s_rho, E = f->eta, X = f->xi, plane = E * X;
    if (E > 0xFFFFFFFFull || X > 0xFFFFFFFFull) return WC_ERR_INVALID;
    const double fill = 9.99999993e+36f;
    // dimensions and variables of Wacomm.cpp:936-1014, in its order
    std::vector<std::pair<std::string, uint64_t>> dims = {{"ocean_time", T}, {"s_rho", S}, {"eta_rho", E},
                                                          {"eta_xi", X}};
    std::vector<Att> gatts;
    if (f->global_attributes)
        for (const char *const *a = f->global_attributes; a[0] && a[1]; a += 2) gatts.push_back(text(a[0], a[1]));
    std::vector<Var> vars;
    vars.push_back(Var{"ocean_time", {0}, NC_DOUBLE,
                       {text("long_name", "time since initialization"),
                        text("units", "seconds since 1968-05-23 00:00:00 GMT"), text("calendar", "gregorian"),
                        text("field", "time, scalar, series"), text("_CoordinateAxisType", "Time")},
                       f->ocean_time, 0.0, T});
    vars.push_back(Var{"s_rho", {1}, NC_DOUBLE,
                       {text("long_name", "S-coordinate at RHO-points"), text("positive", "up"),
                        text("standard_name", "ocean_sigma_coordinates"), text("field", "s_rho, scalar"),
                        text("_CoordinateTransformType", "Vertical"), text("_CoordinateAxes", "s_rho"),
                        text("_CoordinateAxisType", "GeoZ"), text("_CoordinateZisPositive", "up")},
                       f->s_rho_values, 0.0, S});
    vars.push_back(Var{"lat_rho", {2, 3}, NC_DOUBLE,
                       {text("long_name", "latitude of rho-points"), text("unit", "degree_north"),
                        text("standard_name", "latitude"), text("field", "lat_rho, scalar"),
                        text("_coordinateaxistype", "lat")},
                       f->lat_rho, 0.0, plane});
    vars.push_back(Var{"lon_rho", {2, 3}, NC_DOUBLE,
                       {text("long_name", "longitude of rho-points"), text("unit", "degree_east"),
                        text("standard_name", "longitude"), text("field", "lon_rho, scalar"),
                        text("_coordinateaxistype", "lon")},
                       f->lon_rho, 0.0, plane});
    vars.push_back(Var{"mask_rho", {2, 3}, NC_DOUBLE,
                       {text("long_name", "mask on RHO-points"), text("coordinates", "lon_rho lat_rho"),
                        text("units", "1")},
                       f->mask_rho, 0.0, plane});
    vars.push_back(Var{"h", {2, 3}, NC_DOUBLE,
                       {text("long_name", "bathymetry at RHO-point"), text("units", "meter"),
                        text("coordinates", "lon_rho lat_rho"), text("field", "bath, scalar")},
                       f->h, 0.0, plane});
    vars.push_back(Var{"zeta", {0, 2, 3}, NC_FLOAT,
                       {text("long_name", "free-surface"), text("units", "meter"), text("time", "ocean_time"),
                        text("coordinates", "lon_rho lat_rho ocean_time"),
                        text("field", "free-surface, scalar, series"), f32("_FillValue", (float)fill)},
                       f->zeta, fill, T * plane});
    vars.push_back(Var{"conc", {0, 1, 2, 3}, NC_FLOAT,
                       {text("long_name", "concentration_of_suspended_matter_in_sea_water"), text("units", "1"),
                        text("coordinates", "lon_rho lat_rho s_rho ocean_time"), text("field", ""),
                        text("time", "ocean_time"), f32("_FillValue", (float)fill)},
                       f->conc, fill, T * S * plane});
    return write_file(path, 2, dims, gatts, vars);
}

extern "C" int wc_write_restart_nc(const char *path, const wc_restart_file *f) {
    if (!path || !f || !f->ocean_model) return WC_ERR_INVALID;
    const uint64_t n = f->n;
    if (n && (!f->id || !f->lat || !f->lon || !f->depth || !f->k || !f->j || !f->i || !f->health || !f->age ||
              !f->time))
        return WC_ERR_INVALID;
    static const char *const fixed[] = {"wacomm_restart_version", "tracking_direction", "ocean_model",
                                        "checkpoint_time", "random_seed"};
    // attributes in the order of Particles.cpp:154-160 (provenance first)
    std::vector<Att> gatts;
    bool scheme = false;
    if (f->global_attributes)
        for (const char *const *a = f->global_attributes; a[0] && a[1]; a += 2) {
            for (const char *x : fixed)
                if (!std::strcmp(a[0], x)) return WC_ERR_INVALID;
            for (const Att &g : gatts)
                if (g.name == a[0]) return WC_ERR_INVALID;
            scheme = scheme || !std::strcmp(a[0], "wacomm_random_key_scheme");
            gatts.push_back(text(a[0], a[1]));
        }
    if (!scheme) gatts.push_back(text("wacomm_random_key_scheme", kRandomKeyScheme));
    gatts.push_back(text("wacomm_restart_version", "3"));
    gatts.push_back(text("tracking_direction", f->backward ? "backward" : "forward"));
    gatts.push_back(text("ocean_model", f->ocean_model));
    gatts.push_back(f64("checkpoint_time", f->particle_time));
    gatts.push_back(u64("random_seed", f->random_seed));
    // an empty history keeps zero-length variables: particles is then the
    // record dimension with no records
    std::vector<std::pair<std::string, uint64_t>> dims = {{"particles", n}, {"particle_time", 1}};
    const Att name_time = text("long_name", "time since initialization");
    const Att units_time = text("units", "seconds since 1968-05-23 00:00:00 GMT");
    const Att calendar = text("calendar", "gregorian");
    const Att coords = text("coordinates", "ocean_time lon lat depth");
    std::vector<Var> vars;
    vars.push_back(Var{"particle_time", {1}, NC_DOUBLE,
                       {name_time, units_time, calendar, text("field", "time, scalar, series"),
                        text("_CoordinateAxisType", "Time")},
                       &f->particle_time, 0.0, 1});
    vars.push_back(Var{"id", {0}, NC_UINT64, {text("long_name", "id"), text("cf_role", "trajectory_id")}, f->id, 0.0,
                       n});
    vars.push_back(Var{"object_type", {0}, NC_UINT, {text("long_name", "drift object type identifier")},
                       f->object_type, 0.0, n});
    vars.push_back(Var{"drift_side", {0}, NC_INT,
                       {text("long_name", "crosswind orientation: -1 left, 0 undefined, 1 right")}, f->drift_side, 0.0,
                       n});
    vars.push_back(Var{"lat", {0, 1}, NC_DOUBLE,
                       {text("long_name", "latitude of particle"), text("unit", "degree_north"),
                        text("standard_name", "latitude")},
                       f->lat, 0.0, n});
    vars.push_back(Var{"lon", {0, 1}, NC_DOUBLE,
                       {text("long_name", "longitude of particle"), text("unit", "degree_east"),
                        text("standard_name", "longitude")},
                       f->lon, 0.0, n});
    vars.push_back(Var{"depth", {0, 1}, NC_DOUBLE,
                       {text("long_name", "particle depth below sea surface"), text("unit", "meters"),
                        text("standard_name", "depth"), text("positive", "up"), text("axis", "Z")},
                       f->depth, 0.0, n});
    vars.push_back(Var{"k", {0, 1}, NC_DOUBLE, {text("long_name", "depth fractional index")}, f->k, 0.0, n});
    vars.push_back(Var{"j", {0, 1}, NC_DOUBLE, {text("long_name", "latitude fractional index")}, f->j, 0.0, n});
    vars.push_back(Var{"i", {0, 1}, NC_DOUBLE, {text("long_name", "longitude fractional index")}, f->i, 0.0, n});
    vars.push_back(Var{"health", {0, 1}, NC_DOUBLE,
                       {text("long_name", "health of the particle"), text("units", "1e-9"), coords}, f->health, 0.0,
                       n});
    vars.push_back(Var{"age", {0, 1}, NC_DOUBLE,
                       {text("long_name", "age of the particle"), text("units", "seconds since emission"), coords},
                       f->age, 0.0, n});
    vars.push_back(Var{"time", {0, 1}, NC_DOUBLE,
                       {text("long_name", "emission time of the particle"), units_time, calendar, coords}, f->time,
                       0.0, n});
    return write_file(path, 5, dims, gatts, vars);
}

extern "C" int wc_read_restart_nc(const char *path, const wc_config_data *cfg, const char *ocean_model,
                                  wc_particle_data *out, size_t capacity, size_t *count, double *checkpoint_time,
                                  char *err, size_t err_len) {
    set_err(err, err_len, "");
    if (!path || !cfg || !ocean_model || !count) return WC_ERR_INVALID;
    *count = 0;
    NcFile f;
    {
        FILE *fp = fopen(path, "rb");
        if (!fp) {
            set_err(err, err_len, std::string("cannot open ") + path);
            return WC_ERR_IO;
        }
        uint8_t buf[1 << 16];
        size_t got;
        while ((got = fread(buf, 1, sizeof buf, fp)) > 0) f.bytes.insert(f.bytes.end(), buf, buf + got);
        fclose(fp);
    }
    std::string why;
    int rc = parse(f, why);
    if (rc) {
        set_err(err, err_len, why);
        return rc;
    }
    auto bad = [&](const std::string &msg) {
        set_err(err, err_len, msg);
        return WC_ERR_INVALID;
    };
    // restart metadata checks (Particles.cpp:378-432), when the file carries a version
    if (const RAtt *ver = f.att("wacomm_restart_version")) {
        const std::string v = att_text(*ver);
        if (v != "2" && v != "3") return bad("Unsupported restart version: " + v);
        if (v == "2" && cfg->driftModel == 1)
            return bad("Restart version 2 does not contain drift object state required by drift.model=leeway");
        const RAtt *dir = f.att("tracking_direction");
        if (!dir) return bad("Restart file has no tracking_direction attribute");
        const std::string rd = att_text(*dir), cd = cfg->trackingDirection == -1 ? "backward" : "forward";
        if (rd != cd) return bad("Restart tracking direction is " + rd + " but configuration direction is " + cd);
        if (const RAtt *cp = f.att("checkpoint_time")) {
            if (!type_size(cp->type) || cp->type == NC_CHAR || cp->n < 1)
                return bad("Restart checkpoint_time is not numeric");
            if (checkpoint_time) *checkpoint_time = read_num(cp->values.data(), cp->type).d;
        }
        if (const RAtt *sd = f.att("random_seed")) {
            if (sd->type == NC_CHAR || sd->n < 1) return bad("Restart random_seed is not numeric");
            const Num s = read_num(sd->values.data(), sd->type);
            if (s.u != cfg->randomSeed || s.neg || !s.integral)
                return bad("Restart random seed does not match the configured random seed");
        }
        if (const RAtt *sc = f.att("wacomm_random_key_scheme")) {
            const std::string s = att_text(*sc);
            if (s != kRandomKeyScheme)
                return bad("Restart random-key scheme is " + s + " but this build uses " + kRandomKeyScheme);
        }
        if (const RAtt *om = f.att("ocean_model")) {
            const std::string m = att_text(*om);
            if (m != ocean_model)
                return bad("Restart ocean model is " + m + " but configuration ocean model is " + ocean_model);
        }
    }
    // variables (Particles.cpp:434-479)
    const char *names[7] = {"id", "k", "j", "i", "health", "age", "time"};
    const RVar *v[7];
    for (int q = 0; q < 7; ++q) {
        v[q] = f.var(names[q]);
        if (!v[q] || v[q]->type == NC_CHAR) return bad(std::string("Restart file has no numeric variable ") + names[q]);
    }
    const uint64_t n = first_dim_size(f, *v[0]);
    for (int q = 1; q < 7; ++q)
        if (first_dim_size(f, *v[q]) < n) return bad(std::string("Restart variable ") + names[q] + " is shorter than id");
    const RVar *vo = f.var("object_type"), *vs = f.var("drift_side");
    if ((vo && (vo->type == NC_CHAR || first_dim_size(f, *vo) < n)) ||
        (vs && (vs->type == NC_CHAR || first_dim_size(f, *vs) < n)))
        return bad("Restart drift variables do not match id");
    // the count comes from the header: every variable's last element must lie
    // in the file before it is reported (offsets grow with the index)
    if (n > 0) {
        uint64_t off;
        for (const RVar *x : {v[0], v[1], v[2], v[3], v[4], v[5], v[6], vo, vs})
            if (x && !element_offset(f, *x, n - 1, off)) {
                set_err(err, err_len, "truncated NetCDF data");
                return WC_ERR_IO;
            }
    }
    *count = (size_t)n;
    if (!out) return WC_OK;
    if (capacity < n) return bad("Restart holds " + std::to_string(n) + " particles, capacity " +
                                 std::to_string(capacity));
    for (uint64_t q = 0; q < n; ++q) {
        uint64_t off;
        Num x[7];
        for (int c = 0; c < 7; ++c) {
            if (!element_offset(f, *v[c], q, off)) {
                set_err(err, err_len, "truncated NetCDF data");
                return WC_ERR_IO;
            }
            x[c] = read_num(f.bytes.data() + off, v[c]->type);
        }
        if (x[0].neg || !x[0].integral) return bad("Restart id out of range");
        wc_particle_data p{};
        p.id = x[0].u;
        p.k = x[1].d; p.j = x[2].d; p.i = x[3].d; p.health = x[4].d; p.age = x[5].d; p.time = x[6].d;
        if (vo) {
            if (!element_offset(f, *vo, q, off)) { set_err(err, err_len, "truncated NetCDF data"); return WC_ERR_IO; }
            p.driftObjectType = (uint16_t)(unsigned int)read_num(f.bytes.data() + off, vo->type).u;
        }
        if (vs) {
            if (!element_offset(f, *vs, q, off)) { set_err(err, err_len, "truncated NetCDF data"); return WC_ERR_IO; }
            p.driftSide = (int8_t)(int)(int64_t)read_num(f.bytes.data() + off, vs->type).u;
        }
        out[q] = p;
    }
    return WC_OK;
}
