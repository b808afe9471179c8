// Concentration output in the NetCDF classic format (host code).
//
// Wacomm::save (Wacomm.cpp:913-1020) writes, through the NetCDF library, the
// window's time axis, s_rho, lat/lon/mask/h on the rho grid, zeta and the
// concentration frames conc[ocean_time][s_rho][eta_rho][eta_xi] with their
// attributes and a _FillValue of 9.99999993e36.  This writer produces the
// same dimensions, variables and attributes in the NetCDF classic format with
// 64-bit offsets (CDF-2, NetCDF Classic Format Specification: magic "CDF\2",
// big-endian header of dimension, attribute and variable lists, each
// variable's data at its begin offset, 4-byte aligned), readable by every
// NetCDF reader without the library.  The file is written as <path>.partial
// and renamed when complete, so a reader never sees a partial file (the
// reference's output-publication rule).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wacomm_hip.h"

namespace {

enum { NC_CHAR = 2, NC_FLOAT = 5, NC_DOUBLE = 6 };
constexpr uint32_t kDimTag = 0x0A, kVarTag = 0x0B, kAttTag = 0x0C;

struct Buf {
    std::vector<uint8_t> b;
    void u32(uint32_t v) {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void u64(uint64_t v) {
        for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
    void pad() {
        while (b.size() % 4) b.push_back(0);
    }
    void name(const std::string &s) {
        u32((uint32_t)s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
};

struct Att {
    std::string name;
    int type;
    std::vector<uint8_t> payload;   // big-endian values
    uint32_t n;
};

Att text(const std::string &name, const std::string &v) {
    return Att{name, NC_CHAR, std::vector<uint8_t>(v.begin(), v.end()), (uint32_t)v.size()};
}
Att f32(const std::string &name, float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    Buf b;
    b.u32(u);
    return Att{name, NC_FLOAT, b.b, 1};
}

struct Var {
    std::string name;
    std::vector<int> dims;
    int type;
    std::vector<Att> atts;
    const void *data;      // host values in native order, or nullptr (written as `fill`)
    float fill;
    uint64_t count;
    uint64_t begin = 0;
};

void put_atts(Buf &h, const std::vector<Att> &atts) {
    if (atts.empty()) {
        h.u32(0);
        h.u32(0);
        return;
    }
    h.u32(kAttTag);
    h.u32((uint32_t)atts.size());
    for (const Att &a : atts) {
        h.name(a.name);
        h.u32((uint32_t)a.type);
        h.u32(a.n);
        h.b.insert(h.b.end(), a.payload.begin(), a.payload.end());
        h.pad();
    }
}

Buf header(const std::vector<std::pair<std::string, uint64_t>> &dims, const std::vector<Att> &gatts,
           const std::vector<Var> &vars) {
    Buf h;
    h.b = {'C', 'D', 'F', 2};
    h.u32(0);                                   // numrecs: no record dimension
    h.u32(kDimTag);
    h.u32((uint32_t)dims.size());
    for (const auto &d : dims) {
        h.name(d.first);
        h.u32((uint32_t)d.second);
    }
    put_atts(h, gatts);
    h.u32(kVarTag);
    h.u32((uint32_t)vars.size());
    for (const Var &v : vars) {
        h.name(v.name);
        h.u32((uint32_t)v.dims.size());
        for (int d : v.dims) h.u32((uint32_t)d);
        put_atts(h, v.atts);
        h.u32((uint32_t)v.type);
        const uint64_t bytes = v.count * (v.type == NC_DOUBLE ? 8 : 4);
        const uint64_t vsize = (bytes + 3) / 4 * 4;
        h.u32(vsize > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)vsize);
        h.u64(v.begin);
    }
    return h;
}

bool write_data(FILE *fp, const Var &v) {
    const size_t chunk = 1 << 18;
    std::vector<uint8_t> out;
    const bool dbl = v.type == NC_DOUBLE;
    for (uint64_t q0 = 0; q0 < v.count; q0 += chunk) {
        const uint64_t m = v.count - q0 < chunk ? v.count - q0 : chunk;
        out.resize(m * (dbl ? 8 : 4));
        uint8_t *o = out.data();
        for (uint64_t q = 0; q < m; ++q) {
            if (dbl) {
                uint64_t u;
                const double x = v.data ? ((const double *)v.data)[q0 + q] : (double)v.fill;
                std::memcpy(&u, &x, 8);
                for (int s = 56; s >= 0; s -= 8) *o++ = (uint8_t)(u >> s);
            } else {
                uint32_t u;
                const float x = v.data ? ((const float *)v.data)[q0 + q] : v.fill;
                std::memcpy(&u, &x, 4);
                for (int s = 24; s >= 0; s -= 8) *o++ = (uint8_t)(u >> s);
            }
        }
        if (fwrite(out.data(), 1, out.size(), fp) != out.size()) return false;
    }
    const uint64_t bytes = v.count * (dbl ? 8 : 4);
    static const uint8_t zero[4] = {0, 0, 0, 0};
    return bytes % 4 == 0 || fwrite(zero, 1, 4 - bytes % 4, fp) == 4 - bytes % 4;
}

}  // namespace

extern "C" int wc_write_concentration_nc(const char *path, const wc_conc_file *f) {
    if (!path || !f || f->T < 1 || f->s_rho < 1 || f->eta < 1 || f->xi < 1 || !f->ocean_time || !f->s_rho_values ||
        !f->lat_rho || !f->lon_rho || !f->mask_rho || !f->h)
        return WC_ERR_INVALID;
    const uint64_t T = f->T, S = f->s_rho, E = f->eta, X = f->xi, plane = E * X;
    if (E > 0xFFFFFFFFull || X > 0xFFFFFFFFull) return WC_ERR_INVALID;
    const float fill = 9.99999993e+36f;
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
                       f->ocean_time, 0.0f, T});
    vars.push_back(Var{"s_rho", {1}, NC_DOUBLE,
                       {text("long_name", "S-coordinate at RHO-points"), text("positive", "up"),
                        text("standard_name", "ocean_sigma_coordinates"), text("field", "s_rho, scalar"),
                        text("_CoordinateTransformType", "Vertical"), text("_CoordinateAxes", "s_rho"),
                        text("_CoordinateAxisType", "GeoZ"), text("_CoordinateZisPositive", "up")},
                       f->s_rho_values, 0.0f, S});
    vars.push_back(Var{"lat_rho", {2, 3}, NC_DOUBLE,
                       {text("long_name", "latitude of rho-points"), text("unit", "degree_north"),
                        text("standard_name", "latitude"), text("field", "lat_rho, scalar"),
                        text("_coordinateaxistype", "lat")},
                       f->lat_rho, 0.0f, plane});
    vars.push_back(Var{"lon_rho", {2, 3}, NC_DOUBLE,
                       {text("long_name", "longitude of rho-points"), text("unit", "degree_east"),
                        text("standard_name", "longitude"), text("field", "lon_rho, scalar"),
                        text("_coordinateaxistype", "lon")},
                       f->lon_rho, 0.0f, plane});
    vars.push_back(Var{"mask_rho", {2, 3}, NC_DOUBLE,
                       {text("long_name", "mask on RHO-points"), text("coordinates", "lon_rho lat_rho"),
                        text("units", "1")},
                       f->mask_rho, 0.0f, plane});
    vars.push_back(Var{"h", {2, 3}, NC_DOUBLE,
                       {text("long_name", "bathymetry at RHO-point"), text("units", "meter"),
                        text("coordinates", "lon_rho lat_rho"), text("field", "bath, scalar")},
                       f->h, 0.0f, plane});
    vars.push_back(Var{"zeta", {0, 2, 3}, NC_FLOAT,
                       {text("long_name", "free-surface"), text("units", "meter"), text("time", "ocean_time"),
                        text("coordinates", "lon_rho lat_rho ocean_time"),
                        text("field", "free-surface, scalar, series"), f32("_FillValue", fill)},
                       f->zeta, fill, T * plane});
    vars.push_back(Var{"conc", {0, 1, 2, 3}, NC_FLOAT,
                       {text("long_name", "concentration_of_suspended_matter_in_sea_water"), text("units", "1"),
                        text("coordinates", "lon_rho lat_rho s_rho ocean_time"), text("field", ""),
                        text("time", "ocean_time"), f32("_FillValue", fill)},
                       f->conc, fill, T * S * plane});
    // two passes: the header's size fixes each variable's begin offset
    uint64_t off = header(dims, gatts, vars).b.size();
    for (Var &v : vars) {
        v.begin = off;
        off += (v.count * (v.type == NC_DOUBLE ? 8 : 4) + 3) / 4 * 4;
    }
    const Buf h = header(dims, gatts, vars);
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
