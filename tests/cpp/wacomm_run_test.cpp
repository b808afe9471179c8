// wacomm_hip::Wacomm::run and Concentration::evaluate on the device, through
// the C++ host layer only.  The plain configuration of WaComM++'s CPU
// benchmark (120x100x10 grid at 200 m, uniform u = 0.2, v = 0.1 m/s, no land,
// 1,000 particles at (j, i) = (59.5, 49.5), k = -0.2, midpoint, diffusion
// off), 12 hourly intervals.  Checks:
//   - j after 12 h = 59.5 + v * 43200 / D_y with v the stored float 0.1f:
//     D_y is the same haversine length in every row of a rectilinear grid, so
//     j is linear in time (tests/ApplicationEndToEndTest.cpp:252-257 uses the
//     same closed form).  The reference sums its bilinear weights and float
//     products in float, so each step's v carries ~1e-9 relative rounding:
//     the bar is 1e-7 cells on the 21.6-cell displacement (the FP64 oracle
//     lands 2.1e-8 cells from the closed form, bit-identical to the device);
//   - k unchanged (w = 0), every particle alive, all particles identical;
//   - i between the bounds set by the widest and narrowest D_x rows crossed;
//   - every frame of Conc() counts all 1,000 particles in one cell, and
//     Concentration::evaluate adds the same count into a caller frame.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <iostream>

#include "wacomm_hip.hpp"

using namespace wacomm_hip;

int main() {
    const int s_rho = 10, eta = 120, xi = 100, T = 13;
    const double R = 6371000.0, spacing = 200.0, lat0 = 40.75 * M_PI / 180.0;
    auto ocean = std::make_shared<OceanWindow>();
    OceanWindow &o = *ocean;
    o.s_rho = s_rho; o.eta = eta; o.xi = xi;
    const double dlat = spacing / R, dlon = spacing / (R * std::cos(lat0));
    for (int t = 0; t < T; ++t) o.oceanTime.push_back(3600.0 * t);
    for (int j = 0; j < eta; ++j)
        for (int i = 0; i < xi; ++i) {
            o.lonRad.push_back(14.0 * M_PI / 180.0 + i * dlon);
            o.latRad.push_back(lat0 + j * dlat);
            o.h.push_back(50.0);
            o.mask.push_back(1.0);
        }
    o.depthIntervals.assign(s_rho, 0.1);
    const size_t plane = (size_t)eta * xi;
    o.zeta.assign(T * plane, 0.0f);
    o.u.assign(T * s_rho * plane, 0.2f);
    o.v.assign(T * s_rho * plane, 0.1f);
    o.w.assign(T * (s_rho + 1) * plane, 0.0f);
    o.akt.assign(T * (s_rho + 1) * plane, 0.0f);

    auto particles = std::make_shared<Particles>();
    for (int q = 0; q < 1000; ++q) particles->push_back(Particle(q, -0.2, 59.5, 49.5, 0.0));
    auto config = std::make_shared<wc_config_data>(defaultConfig());
    config->random = false;
    Wacomm wacomm(config, ocean, std::make_shared<Sources>(), particles);
    double time = 0, part = 0, cuda = 0;
    int nParticles = 0, idx = 0;
    wacomm.run(time, part, cuda, nParticles, idx);

    assert(nParticles == 1000 && particles->size() == 1000);
    // the library's haversine cell sizes (Particle.cpp:548-574)
    std::vector<double> dx(plane), dy(plane);
    wc_cell_metric_table(eta, xi, o.lonRad.data(), o.latRad.data(), dx.data(), dy.data());
    const double Dy = dy[59 * xi + 49];
    // the same length in every row up to the haversine's rounding (~4e-12
    // relative), which moves j by < 1e-10 cells over the 21.6 cells crossed
    for (int j = 0; j + 1 < eta; ++j)
        for (int i = 0; i + 1 < xi; ++i) assert(std::fabs(dy[(size_t)j * xi + i] - Dy) < 1e-11 * Dy);
    const double v = 0.1f, u = 0.2f;                 // the stored float currents
    const double jEnd = 59.5 + v * 43200.0 / Dy;
    const Particle &p0 = (*particles)[0];
    std::cout.precision(17);
    std::cout << "j=" << p0.J() << " expected " << jEnd << "  i=" << p0.I() << " k=" << p0.K() << std::endl;
    assert(std::fabs(p0.J() - jEnd) < 1e-7);
    assert(p0.K() == -0.2);
    double DxMin = 1e30, DxMax = 0;
    for (int j = 59; j <= (int)jEnd; ++j) {
        DxMin = std::fmin(DxMin, dx[(size_t)j * xi + 49]);
        DxMax = std::fmax(DxMax, dx[(size_t)j * xi + 49]);
    }
    assert(p0.I() > 49.5 + u * 43200.0 / DxMax - 1e-7 && p0.I() < 49.5 + u * 43200.0 / DxMin + 1e-7);
    for (const Particle &p : *particles) {
        assert(p.isAlive());
        assert(p.K() == p0.K() && p.J() == p0.J() && p.I() == p0.I() && p.Age() == 43200.0);
    }
    assert(cuda > 0 && time > 0);
    assert(wacomm.ParticleSteps() == 1000ull * 12 * 120);

    // frames 1..12 each hold the 1,000 particles in one cell
    const size_t frame = (size_t)s_rho * plane;
    const std::vector<float> &conc = wacomm.Conc();
    for (int t = 1; t < T; ++t) {
        double total = 0, top = 0;
        for (size_t q = 0; q < frame; ++q) {
            total += conc[t * frame + q];
            top = std::fmax(top, conc[t * frame + q]);
        }
        assert(total == 1000 && top == 1000);
    }
    // Concentration::evaluate into a caller frame holding 5 in the target cell
    std::vector<float> mine(2 * frame, 0.0f);
    const int kr = (int)std::round(p0.K()), jr = (int)std::round(p0.J()), ir = (int)std::round(p0.I());
    const size_t cell = (size_t)(kr + s_rho - 1) * plane + (size_t)jr * xi + ir;
    mine[frame + cell] = 5.0f;
    Context &ctx = wacomm.context();
    Concentration::evaluate(ctx, particles.get(), 1, s_rho, eta, xi, mine.data());
    assert(mine[frame + cell] == 1005.0f);
    double total = 0;
    for (float x : mine) total += x;
    assert(total == 1005.0);
    // Wacomm::save: the concentration file exists, complete (no .partial left)
    wacomm.save("wc-conc-test.nc");
    {
        std::FILE *fp = std::fopen("wc-conc-test.nc", "rb");
        assert(fp);
        char magic[4];
        assert(std::fread(magic, 1, 4, fp) == 4 && magic[0] == 'C' && magic[1] == 'D' && magic[2] == 'F' &&
               magic[3] == 2);
        std::fseek(fp, 0, SEEK_END);
        assert((size_t)std::ftell(fp) > (size_t)T * s_rho * plane * 4);
        std::fclose(fp);
        assert(std::fopen("wc-conc-test.nc.partial", "rb") == nullptr);
        std::remove("wc-conc-test.nc");
    }
    std::cout << "wacomm_run_test: PASS" << std::endl;
    return 0;
}
