// NetCDF restart of wacomm_hip::Particles: Particles::loadFromNetCDF
// (Particles.cpp:373-480) on files written through the C ABI
// (wc_write_restart_nc, the layout of Particles.cpp:150-264).  CPU only: no
// device call is made.  Checks the exact 64-bit identity and full double
// precision of tests/RestartTest.cpp:104-122, the drift state, the
// checkpoint_time -> restartCheckpoint hand-over, and that a mismatched
// tracking direction or random seed throws std::runtime_error with the
// reference's message.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "wacomm_hip.hpp"

using namespace wacomm_hip;

static void write(const std::string &name, int backward, std::uint64_t seed) {
    const std::uint64_t id[2] = {9007199254740993ULL, 18446744073709551615ULL};
    const double k[2] = {-.123456789012345, -0.0}, j[2] = {5.85472728621506, 5e-324},
                 i[2] = {5.32375564640966, 1.7976931348623157e308}, health[2] = {.920044414487148, 1.0},
                 age[2] = {7200.0000133514404, 0.0}, time[2] = {1675900800.25, -1.0}, lat[2] = {40.1, 40.2},
                 lon[2] = {14.1, 14.2}, depth[2] = {-1.0, -2.0};
    const std::uint32_t object[2] = {1, 0};
    const std::int32_t side[2] = {1, -1};
    const char *prov[] = {"wacomm_git_revision", "test", nullptr, nullptr};
    wc_restart_file f{2, 3600.5, seed, backward, "roms", prov, id, object, side,
                      lat, lon, depth, k, j, i, health, age, time};
    const int rc = wc_write_restart_nc(name.c_str(), &f);
    assert(rc == WC_OK);
}

static std::string rejection(const std::string &name, wc_config_data cfg) {
    try {
        Particles p;
        p.loadFromNetCDF(name, cfg, "roms");
    } catch (const std::runtime_error &e) {
        return e.what();
    }
    return "";
}

int main() {
    const std::string file = "wc-restart-test.nc";
    write(file, 0, 5489);
    wc_config_data cfg = defaultConfig();
    Particles read;
    read.loadFromNetCDF(file, cfg, "roms");
    assert(read.size() == 2);
    assert(read[0].Id() == 9007199254740993ULL && read[1].Id() == 18446744073709551615ULL);
    assert(read[0].K() == -.123456789012345 && read[0].J() == 5.85472728621506 && read[0].I() == 5.32375564640966);
    assert(read[0].Health() == .920044414487148 && read[0].Age() == 7200.0000133514404);
    assert(read[0].Time() == 1675900800.25);
    assert(std::signbit(read[1].K()) && read[1].J() == 5e-324 && read[1].I() == 1.7976931348623157e308);
    assert(read[0].DriftObject() == 1 && read[0].Side() == 1 && read[1].DriftObject() == 0 && read[1].Side() == -1);
    assert(cfg.restartCheckpoint == 3600.5);

    wc_config_data backward = defaultConfig();
    backward.trackingDirection = -1;
    std::string msg = rejection(file, backward);
    assert(msg == "Restart tracking direction is forward but configuration direction is backward");
    wc_config_data seed = defaultConfig();
    seed.randomSeed = 7;
    msg = rejection(file, seed);
    assert(msg == "Restart random seed does not match the configured random seed");
    try {
        Particles p;
        wc_config_data c = defaultConfig();
        p.loadFromNetCDF(file, c, "nemo");
        assert(false);
    } catch (const std::runtime_error &e) {
        assert(std::string(e.what()) == "Restart ocean model is roms but configuration ocean model is nemo");
    }
    std::remove(file.c_str());
    std::cout << "PASS particles_nc_test" << std::endl;
    return 0;
}
