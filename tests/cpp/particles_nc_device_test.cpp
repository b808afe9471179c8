// Particles::saveAsNetCDF (Particles.cpp:150-264) through the C++ layer: the
// records (netcdfRecords with kji2deplatlon, OceanModelAdapter.cpp:271-321)
// are computed on the device, written, and read back with loadFromNetCDF.
// Grid: 6x5 rectilinear, lat 40 + 0.1 j, lon 10 + 0.1 i (degrees), h = 80 +
// 20 i + 10 j, five W levels -1, -0.6, -0.3, -0.1, 0 (the fixture of
// tests/VerticalCoordinateTest.cpp:73-90, widened).  Checks: only the
// in-domain particles are kept, in order, with their drift state; lat and lon
// are linear in j and i (1e-12); depth at an integer level L is
// -h |sW(L)| (1e-9).
#include <cassert>
#include <cmath>
#include <cstdio>
#include <iostream>

#include "wacomm_hip.hpp"

using namespace wacomm_hip;

int main() {
    const int eta = 6, xi = 5, s_rho = 4, s_w = 5;
    const double sW[s_w] = {-1.0, -0.6, -0.3, -0.1, 0.0};
    std::vector<double> lat, lon, latRad, lonRad, h, mask(eta * xi, 1.0), di;
    for (int j = 0; j < eta; ++j)
        for (int i = 0; i < xi; ++i) {
            lat.push_back(40 + 0.1 * j);
            lon.push_back(10 + 0.1 * i);
            latRad.push_back(0.0174533 * lat.back());
            lonRad.push_back(0.0174533 * lon.back());
            h.push_back(80 + 20 * i + 10 * j);
        }
    for (int k = 1; k < s_w; ++k) di.push_back(sW[k] - sW[k - 1]);
    Context ctx(0);
    ctx.check(wc_set_grid(ctx.get(), s_w, s_rho, eta, xi, lonRad.data(), latRad.data(), h.data(), mask.data(),
                          di.data(), 2, WC_STORE_F32),
              "wc_set_grid");
    ctx.check(wc_set_geo(ctx.get(), lat.data(), lon.data(), sW), "wc_set_geo");

    Particles ps;
    ps.push_back(Particle(10, -1.0, 1.25, 2.5, 0.9, 100.0, 0.0));     // kept
    ps.push_back(Particle(11, -0.5, -0.5, 2.0, 1.0, 0.0, 0.0));      // j < 0
    ps.push_back(Particle(12, -3.0, 4.75, 3.5, 0.5, 10.0, 5.0));     // kept, lowest valid level
    ps.push_back(Particle(13, -0.5, 5.0, 1.0, 1.0, 0.0, 0.0));       // north node line
    ps.push_back(Particle(14, -4.0, 1.0, 1.0, 1.0, 0.0, 0.0));       // below the levels
    ps.push_back(Particle(15, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0));        // kept, surface, corner
    ps[0].Drift(3, -1);
    ps[5].Drift(2, 1);
    wc_config_data cfg = defaultConfig();
    const std::string file = "wc-device-restart.nc";
    ps.saveAsNetCDF(ctx, file, 7200.0, cfg, "roms");

    Particles back;
    back.loadFromNetCDF(file, cfg, "roms");
    assert(back.size() == 3);
    const int kept[3] = {0, 2, 5};
    for (int r = 0; r < 3; ++r) {
        const Particle &a = back[r], &b = ps[kept[r]];
        assert(a.Id() == b.Id() && a.K() == b.K() && a.J() == b.J() && a.I() == b.I());
        assert(a.Health() == b.Health() && a.Age() == b.Age() && a.Time() == b.Time());
        assert(a.DriftObject() == b.DriftObject() && a.Side() == b.Side());
    }
    assert(cfg.restartCheckpoint == 7200.0);

    // the records themselves (lat, lon, depth, drift state) from the device
    size_t n = 0;
    ctx.check(wc_set_particles(ctx.get(), ps.raw(), ps.size()), "wc_set_particles");
    std::vector<std::uint64_t> id(6);
    std::vector<double> la(6), lo(6), dep(6), k(6), j(6), i(6), he(6), ag(6), ti(6);
    std::vector<std::uint32_t> ot(6);
    std::vector<std::int32_t> sd(6);
    wc_records rec{id.data(), la.data(), lo.data(), dep.data(), k.data(), j.data(), i.data(), he.data(),
                   ag.data(), ti.data(), ot.data(), sd.data()};
    ctx.check(wc_particle_records(ctx.get(), &rec, 6, &n), "wc_particle_records");
    assert(n == 3);
    assert(ot[0] == 3 && sd[0] == -1 && ot[1] == 0 && sd[1] == 0 && ot[2] == 2 && sd[2] == 1);
    for (size_t r = 0; r < n; ++r) {
        assert(std::fabs(la[r] - (40 + 0.1 * j[r])) < 1e-12 && std::fabs(lo[r] - (10 + 0.1 * i[r])) < 1e-12);
        const double hh = 80 + 20 * i[r] + 10 * j[r];
        assert(std::fabs(dep[r] + hh * std::fabs(sW[(int)k[r] + s_w - 1])) < 1e-9);
    }
    std::remove(file.c_str());
    std::cout << "PASS particles_nc_device_test" << std::endl;
    return 0;
}
