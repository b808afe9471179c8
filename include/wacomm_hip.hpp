/*
 * wacomm_hip.hpp -- C++17 host layer over the C ABI of libwacomm_hip.so with
 * the roles and names of WaComM++'s own host classes, so a C++ host can swap
 * its move/gridding backend for this one:
 *
 *   wacomm_hip::Particle        particle_data wrapper      (Particle.hpp:43-97)
 *   wacomm_hip::Particles       vector of particles, text restarts
 *                                                          (Particles.hpp:17-56,
 *                                                           Particles.cpp:28-55, 320-364)
 *   wacomm_hip::OceanWindow     the forcing window         (oceanmodel_data,
 *                                                           OceanModelAdapter.hpp:20-34)
 *   wacomm_hip::Sources         release points + schedules (Sources.hpp, SourceEmission.hpp)
 *   wacomm_hip::Wacomm          Wacomm::run's interval loop (Wacomm.hpp:30-41,
 *                                                           Wacomm.cpp:89-770)
 *   wacomm_hip::Concentration::evaluate                    (Concentration.hpp:16-44)
 *
 * Every device call goes through include/wacomm_hip.h; a failing call throws
 * wacomm_hip::Error carrying the wc_status and the library's message.  There
 * is no CPU fallback: without a HIP device the context constructor throws.
 * Header-only; link with -lwacomm_hip.  No HIP headers are needed.
 */
#ifndef WACOMM_HIP_HPP
#define WACOMM_HIP_HPP

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iterator>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "wacomm_hip.h"

namespace wacomm_hip {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

/* The Config::setDefault values that the move reads (Config.cpp:46-202). */
inline wc_config_data defaultConfig() {
    wc_config_data c{};
    c.random = true;
    c.randomSources = true;
    c.deltat = 3600.0;
    c.dti = 30.0;
    c.survprob = 1.0e-4;
    c.tau0 = 86400.0;
    c.sv = 0.0;
    c.crid = 1.0;
    c.sigma = 3.46;
    c.shoreLimit = 0.25;
    c.upperClosure = 1;        // constraint
    c.lowerClosure = 2;        // kill
    c.horizontalClosure = 3;   // reflection
    c.randomSeed = 5489;
    c.trackingDirection = 1;
    c.backwardDiffusion = 0;
    c.integrator = 2;          // midpoint
    c.restartCheckpoint = std::numeric_limits<double>::quiet_NaN();
    c.driftModel = 0;
    return c;
}

/* One device context (wc_ctx), owned. */
class Context {
public:
    explicit Context(int device = 0) {
        int rc = wc_ctx_create(device, &ctx_);
        if (rc != WC_OK) throw Error(rc, "wc_ctx_create(device=" + std::to_string(device) + ") failed");
    }
    ~Context() {
        if (ctx_) wc_ctx_destroy(ctx_);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    wc_ctx *get() const { return ctx_; }
    void check(int rc, const char *call) const {
        if (rc != WC_OK) throw Error(rc, std::string(call) + ": " + wc_last_error(ctx_));
    }

private:
    wc_ctx *ctx_ = nullptr;
};

/* particle_data with the accessors of Particle.hpp:43-97. */
class Particle {
public:
    Particle() = default;
    Particle(std::uint64_t id, double k, double j, double i, double health, double age, double time) {
        d_.id = id; d_.k = k; d_.j = j; d_.i = i; d_.health = health; d_.age = age; d_.time = time;
    }
    Particle(std::uint64_t id, double k, double j, double i, double time) : Particle(id, k, j, i, 1.0, 0.0, time) {}
    explicit Particle(const wc_particle_data &d) : d_(d) {}

    bool isAlive() const { return d_.health > 0; }        // Particle.cpp:65-67
    double K() const { return d_.k; }
    double J() const { return d_.j; }
    double I() const { return d_.i; }
    double Health() const { return d_.health; }
    double Age() const { return d_.age; }
    double Time() const { return d_.time; }
    std::uint64_t Id() const { return d_.id; }
    int DriftObject() const { return d_.driftObjectType; }
    int Side() const { return d_.driftSide; }
    void Drift(int objectType, int side) {
        d_.driftObjectType = (uint16_t)objectType;
        d_.driftSide = (int8_t)side;
    }
    const wc_particle_data &data() const { return d_; }
    void data(const wc_particle_data &d) { d_ = d; }

private:
    wc_particle_data d_{};
};

static_assert(sizeof(Particle) == sizeof(wc_particle_data), "Particle must stay a particle_data");

/* Particles.hpp:17-56: a vector of particles and its text restart format. */
class Particles : private std::vector<Particle> {
public:
    using std::vector<Particle>::push_back;
    using std::vector<Particle>::operator[];
    using std::vector<Particle>::size;
    using std::vector<Particle>::at;
    using std::vector<Particle>::empty;
    using std::vector<Particle>::begin;
    using std::vector<Particle>::end;
    using std::vector<Particle>::erase;
    using std::vector<Particle>::clear;
    using std::vector<Particle>::resize;

    wc_particle_data *raw() { return reinterpret_cast<wc_particle_data *>(data()); }
    const wc_particle_data *raw() const { return reinterpret_cast<const wc_particle_data *>(data()); }

    /* Particles::saveAsTxt (Particles.cpp:28-55): a tab and the count on the
       first line, then "id i j k health age time object side" per particle
       at max_digits10. */
    void saveAsTxt(const std::string &fileName) const {
        std::ofstream out(fileName);
        out << "\t" << (int)size() << "\n";
        out << std::setprecision(std::numeric_limits<double>::max_digits10);
        for (const Particle &p : *this)
            out << p.Id() << " " << p.I() << " " << p.J() << " " << p.K() << " " << p.Health() << " " << p.Age()
                << " " << p.Time() << " " << p.DriftObject() << " " << p.Side() << "\n";
        out.close();
        if (!out) throw std::runtime_error("Unable to write the text history " + fileName);
    }

    /* Particles::loadFromTxt (Particles.cpp:320-364): the first line is
       skipped; 5 tokens are a Fortran WaComM record "i j k health age" (id =
       line number, time = -1); 7 or more are "id i j k health age time
       [object side]"; anything else throws.  As in the reference, a file
       that cannot be opened adds no particles. */
    void loadFromTxt(const std::string &fileName) {
        std::ifstream in(fileName);
        unsigned long count = 0;
        std::string line;
        while (std::getline(in, line)) {
            if (count > 0) {
                std::istringstream iss(line);
                std::vector<std::string> t{std::istream_iterator<std::string>{iss},
                                           std::istream_iterator<std::string>{}};
                std::uint64_t id = count;
                double k, j, i, health, age, time = -1;
                if (t.size() == 5) {
                    i = std::atof(t[0].c_str()); j = std::atof(t[1].c_str()); k = std::atof(t[2].c_str());
                    health = std::atof(t[3].c_str()); age = std::atof(t[4].c_str());
                } else if (t.size() >= 7) {
                    id = std::stoull(t[0]);
                    i = std::atof(t[1].c_str()); j = std::atof(t[2].c_str()); k = std::atof(t[3].c_str());
                    health = std::atof(t[4].c_str()); age = std::atof(t[5].c_str()); time = std::atof(t[6].c_str());
                } else {
                    throw std::runtime_error("Invalid particle restart record");
                }
                Particle p(id, k, j, i, health, age, time);
                if (t.size() >= 9) p.Drift(std::atoi(t[7].c_str()), std::atoi(t[8].c_str()));
                push_back(p);
            }
            count++;
        }
    }

    /* Particles::saveAsNetCDF (Particles.cpp:150-264): restart version 3 of
       the particles inside the interpolation domain (netcdfRecords, with
       depth, lat and lon from kji2deplatlon computed on the device), in the
       NetCDF classic format with 64-bit data.  ctx holds the grid and its
       geographic fields (wc_set_roms_grid, or wc_set_grid + wc_set_geo); its
       population is replaced by these particles.  provenance: NULL or a
       NULL-terminated list of name, value strings. */
    void saveAsNetCDF(Context &ctx, const std::string &fileName, double particleTime, const wc_config_data &config,
                      const std::string &oceanModel, const char *const *provenance = nullptr) const {
        wc_ctx *c = ctx.get();
        ctx.check(wc_set_particles(c, raw(), size()), "wc_set_particles");
        size_t n = 0;
        ctx.check(wc_particle_records(c, nullptr, 0, &n), "wc_particle_records");
        std::vector<std::uint64_t> id(n);
        std::vector<double> f[9];
        for (auto &v : f) v.resize(n);
        std::vector<std::uint32_t> object(n);
        std::vector<std::int32_t> side(n);
        wc_records r{id.data(), f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(),
                     f[5].data(), f[6].data(), f[7].data(), f[8].data(), object.data(), side.data()};
        size_t m = 0;
        ctx.check(wc_particle_records(c, &r, n, &m), "wc_particle_records");
        wc_restart_file file{n, particleTime, config.randomSeed, config.trackingDirection == -1 ? 1 : 0,
                             oceanModel.c_str(), provenance, id.data(), object.data(), side.data(),
                             f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(),
                             f[6].data(), f[7].data(), f[8].data()};
        const int rc = wc_write_restart_nc(fileName.c_str(), &file);
        if (rc != WC_OK) throw Error(rc, "cannot write " + fileName);
    }

    /* Particles::loadFromNetCDF (Particles.cpp:373-480): appends the file's
       particles after the reference's checks of its restart metadata against
       config and oceanModel (std::runtime_error with the reference's message);
       a checkpoint_time in the file becomes config.restartCheckpoint. */
    void loadFromNetCDF(const std::string &fileName, wc_config_data &config, const std::string &oceanModel) {
        size_t n = 0;
        double checkpoint = config.restartCheckpoint;
        char err[512];
        int rc = wc_read_restart_nc(fileName.c_str(), &config, oceanModel.c_str(), nullptr, 0, &n, &checkpoint,
                                    err, sizeof err);
        std::vector<wc_particle_data> p(n);
        if (rc == WC_OK && n)
            rc = wc_read_restart_nc(fileName.c_str(), &config, oceanModel.c_str(), p.data(), n, &n, &checkpoint,
                                    err, sizeof err);
        if (rc == WC_ERR_INVALID) throw std::runtime_error(err);
        if (rc != WC_OK) throw Error(rc, err);
        config.restartCheckpoint = checkpoint;
        for (const wc_particle_data &d : p) push_back(Particle(d));
    }

};

/* The forcing window (oceanmodel_data, OceanModelAdapter.hpp:20-34), host
   side, in the layouts of include/wacomm_hip.h. */
struct OceanWindow {
    int s_rho = 0, eta = 0, xi = 0;
    std::vector<double> oceanTime;               // [T]
    std::vector<double> lonRad, latRad, h, mask; // [eta][xi]
    std::vector<double> depthIntervals;          // [s_w-1], entry kI+s_w-2
    std::vector<float> zeta;                     // [T][eta][xi]
    std::vector<float> u, v;                     // [T][s_rho][eta][xi]
    std::vector<float> w, akt;                   // [T][s_rho+1][eta][xi]
    std::vector<double> sRho;                    // [s_rho] sigma at rho levels (for the output file)
    std::vector<double> lon, lat;                // [eta][xi] degrees (empty: lonRad / 0.0174533)
    int T() const { return (int)oceanTime.size(); }
    int s_w() const { return s_rho + 1; }
};

/* Release points and schedules (Sources.hpp, SourceEmission.hpp:12-16). */
using Sources = std::vector<wc_source>;

/* Wacomm (Wacomm.hpp:30-41): owns a device context and runs the interval
   loop of Wacomm::run over the window on the device. */
class Wacomm {
public:
    Wacomm(std::shared_ptr<wc_config_data> config, std::shared_ptr<OceanWindow> ocean,
           std::shared_ptr<Sources> sources, std::shared_ptr<Particles> particles, int device = 0,
           int path = WC_PATH_PARITY, bool maskOutput = false)
        : config_(std::move(config)), ocean_(std::move(ocean)), sources_(std::move(sources)),
          particles_(std::move(particles)), ctx_(device), path_(path), maskOutput_(maskOutput) {}

    /* Wacomm::run(time, part, cuda, nParticles, idx) (Wacomm.cpp:89-770):
       time = wall ms of the interval loop, part = particles per second,
       cuda = device ms of the move kernels, nParticles = particles alive at
       the end, idx = the last record reached.  The particles are moved in
       place; Conc() holds the [T][s_rho][eta][xi] frames (Wacomm.cpp:172-186,
       land = 1e37 when maskOutput).  Returns 0. */
    int run(double &time, double &part, double &cuda, int &nParticles, int &idx) {
        const OceanWindow &o = *ocean_;
        wc_ctx *c = ctx_.get();
        // three forcing slots: the record after next uploads while the move runs
        ctx_.check(wc_set_grid(c, o.s_w(), o.s_rho, o.eta, o.xi, o.lonRad.data(), o.latRad.data(), o.h.data(),
                               o.mask.empty() ? nullptr : o.mask.data(), o.depthIntervals.data(), 3, WC_STORE_F32),
                   "wc_set_grid");
        ctx_.check(wc_set_particles(c, particles_->raw(), particles_->size()), "wc_set_particles");
        wc_window win{o.T(), o.oceanTime.data(), o.zeta.data(), o.u.data(), o.v.data(), o.w.data(), o.akt.data()};
        conc_.assign((size_t)o.T() * o.s_rho * o.eta * o.xi, 0.0f);
        wc_run_stats st{};
        const wc_source *src = sources_ && !sources_->empty() ? sources_->data() : nullptr;
        const int nsrc = sources_ ? (int)sources_->size() : 0;
        ctx_.check(wc_run_window(c, config_.get(), &win, src, nsrc, path_, conc_.data(), maskOutput_ ? 1 : 0, &st),
                   "wc_run_window");
        size_t n = 0;
        ctx_.check(wc_particle_count(c, &n), "wc_particle_count");
        particles_->resize(n);
        ctx_.check(wc_get_particles(c, particles_->raw(), n), "wc_get_particles");
        time = st.solver_s * 1000.0;
        part = st.solver_s > 0 ? (double)n / st.solver_s : 0.0;
        cuda = st.move_ms;
        nParticles = (int)n;
        idx = config_->trackingDirection < 0 ? 0 : o.T() - 1;
        steps_ = st.particle_steps;
        return 0;
    }

    const std::vector<float> &Conc() const { return conc_; }

    /* Wacomm::save (Wacomm.cpp:913-1020): the window's concentration frames
       with its grid and zeta, in NetCDF classic format. */
    void save(const std::string &fileName) const {
        const OceanWindow &o = *ocean_;
        const size_t plane = (size_t)o.eta * o.xi;
        std::vector<double> lon = o.lon, lat = o.lat, sr = o.sRho, mask = o.mask;
        if (lon.size() != plane || lat.size() != plane) {
            lon.resize(plane);
            lat.resize(plane);
            for (size_t q = 0; q < plane; ++q) {
                lon[q] = o.lonRad[q] / 0.0174533;              // the adapter's degree factor
                lat[q] = o.latRad[q] / 0.0174533;
            }
        }
        if ((int)sr.size() != o.s_rho) {
            sr.resize(o.s_rho);
            for (int k = 0; k < o.s_rho; ++k) sr[k] = (k - o.s_rho + 0.5) / o.s_rho;   // uniform sigma
        }
        if (mask.size() != plane) mask.assign(plane, 1.0);
        const char *attrs[] = {"wacomm_random_key_scheme", "ordered-splitmix64-v2", nullptr, nullptr};
        wc_conc_file f{o.T(), o.s_rho, o.eta, o.xi, o.oceanTime.data(), sr.data(), lat.data(), lon.data(),
                       mask.data(), o.h.data(), o.zeta.data(), conc_.data(), attrs};
        const int rc = wc_write_concentration_nc(fileName.c_str(), &f);
        if (rc != WC_OK) throw Error(rc, "cannot write " + fileName);
    }
    std::uint64_t ParticleSteps() const { return steps_; }
    Context &context() { return ctx_; }

private:
    std::shared_ptr<wc_config_data> config_;
    std::shared_ptr<OceanWindow> ocean_;
    std::shared_ptr<Sources> sources_;
    std::shared_ptr<Particles> particles_;
    Context ctx_;
    int path_;
    bool maskOutput_;
    std::vector<float> conc_;
    std::uint64_t steps_ = 0;
};

namespace Concentration {

/* Concentration::evaluate (Concentration.hpp:16-44): adds one to frame
   concentrationTimeIdx of conc ([T][s_rho][eta][xi], host) for every alive
   particle whose rounded k, j, i lie in the grid.  The device counts each
   cell exactly; for an integer-valued frame the count is added as the
   reference's repeated float +1 would add it, which stops at 2^24.  ctx
   must hold a grid of these sizes (wc_set_grid); its particle population
   is replaced. */
inline void evaluate(Context &ctx, const Particles *particles, int concentrationTimeIdx, size_t s_rho,
                     size_t eta_rho, size_t xi_rho, float *conc) {
    wc_ctx *c = ctx.get();
    ctx.check(wc_set_particles(c, particles->raw(), particles->size()), "wc_set_particles");
    const size_t frame = s_rho * eta_rho * xi_rho;
    std::vector<uint32_t> counts(frame);
    ctx.check(wc_grid(c, counts.data(), nullptr, 0), "wc_grid");
    float *dst = conc + (size_t)concentrationTimeIdx * frame;
    const double cap = 16777216.0;                      // 2^24: x + 1.0f == x from here on
    for (size_t q = 0; q < frame; ++q)
        if (counts[q] && dst[q] < cap) dst[q] = (float)std::fmin((double)dst[q] + counts[q], cap);
}

}  // namespace Concentration

}  // namespace wacomm_hip

#endif
