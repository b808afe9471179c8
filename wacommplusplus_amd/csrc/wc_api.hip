// C ABI of libwacomm_hip.so: drop-in entry points and the native context.
// Declarations and contracts: include/wacomm_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wacomm_hip.h"
#include "wc_launch.h"
#include "wc_philox.h"

#pragma clang fp contract(off)

using namespace wc;

// timed kernel groups: 0 move, 1 grid, 2 ROMS normalisation, 3 cell-tile sort
constexpr int kTimedKernels = 5;

struct wc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // static grid
    bool haveGrid = false;
    int s_w = 0, s_rho = 0, eta = 0, xi = 0;
    int store = WC_STORE_F32;
    double *d_h = nullptr, *d_depth = nullptr, *d_dx = nullptr, *d_dy = nullptr, *d_mask = nullptr;
    double2 *d_metric = nullptr;   // {dx, dy} per cell
    float *d_hf = nullptr, *d_md = nullptr, *d_depthf = nullptr;
    void *d_pack = nullptr;    // gather layout of the mixed-precision path

    // forcing record ring: per slot, the FP32 fields as uploaded
    struct Slot {
        float *f[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};      // zeta, u, v, w, akt
        bool filled = false;
        int record = -1;      // window record held by the slot (wc_run_window)
        // Recorded on the compute stream after the last enqueued work that
        // reads or writes the slot; an upload into the slot waits for it on
        // the copy stream.  Rule: every launch that reads a slot records it
        // afterwards (wc_move_interval, the ensemble step, wc_compare_records),
        // so no reader depends on the host having waited for it.
        hipEvent_t released = nullptr;
    };
    std::vector<Slot> slots;

    // record uploads run on their own stream, so the next record's copy
    // overlaps the move already enqueued on `stream`
    hipStream_t copy = nullptr;

    // particles (SoA, double-buffered for compaction); trackEvents: the
    // moves OR closure events into s.ev
    size_t n = 0, cap = 0;
    bool trackEvents = false;
    // cell-tile order of the population (wc_set_sort): perm lists the
    // population's indices tile by tile; the moves and the gridding visit the
    // particles in that order.  perm2 and remap are scratch of the sort and
    // of compaction (which keeps perm valid); sortBins: tiles + 1 counters.
    uint32_t *d_perm = nullptr, *d_perm2 = nullptr, *d_remap = nullptr, *d_sortBins = nullptr;
    uint32_t *d_scanSums = nullptr;       // kScanSumsCap entries: scratch of the multi-block scans
    int sortTile = 0, sortPeriod = 1;
    long sortCount = 0;
    // LDS staging of the forcing for the mixed-precision move (wc_set_lds)
    LdsTiling lds{0, 0};
    bool permValid = false;
    SoA s{}, s2{};
    uint32_t *d_blockCounts = nullptr;
    unsigned long long *d_total = nullptr;
    unsigned long long *d_steps = nullptr;
    uint32_t *d_counts = nullptr;
    float *d_conc = nullptr;

    // ROMS normalisation (wc_set_roms_grid): per rho point the wet/face bits,
    // {cos, sin} of the grid angle, staging for one record's raw u and v,
    // the error flag
    bool haveRoms = false;
    RomsDims roms{};
    RomsMissing romsMissing{};
    uint8_t *d_face = nullptr;
    double2 *d_rot = nullptr;
    float *d_u_raw = nullptr, *d_v_raw = nullptr;
    uint32_t *d_flag = nullptr;
    // host copy of the ROMS static grid, for the adjacent-file check
    std::vector<double> romsLon, romsLat, romsH, romsMask, romsSw, romsSrho;
    // duplicate-record comparison key
    unsigned long long *d_key = nullptr;

    // geographic grid of the particle records (wc_set_geo): lat, lon
    // [eta][xi], s_w values; the records' depth, lat, lon (3 x recCap)
    bool haveGeo = false;
    double *d_lat = nullptr, *d_lon = nullptr, *d_sw = nullptr;
    double *d_rec = nullptr;
    size_t recCap = 0;

    // batched ensemble (wc_ens_*): compact state, a saved copy for the
    // autotuner, scenario sigmas, per-scenario surface counts, the device
    // step array and counter read by the captured step, the graph
    struct Ens {
        bool have = false;
        size_t n = 0;
        int nScen = 0;
        EnsSoA s{}, save{}, tmp{};
        uint32_t *d_bins = nullptr;
        int sortTile = 0, sortPeriod = 1;
        int ldsChunk = 0, ldsHalo = 2;
        float *d_sigma = nullptr;
        uint32_t *d_surface = nullptr;
        EnsStep *d_steps = nullptr;
        int *d_stepIdx = nullptr;
        unsigned long long *d_count = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int block = 256, ppt = 1, occ = 0;
        bool tuned = false;
        bool graphSurface = false;
    } ens;

    // timing: an event pair per timed launch, resolved when queried (no
    // synchronisation between launches)
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[kTimedKernels];
    std::vector<hipEvent_t> spare;
    hipEvent_t cur = nullptr;
    double ms[kTimedKernels] = {};
    long launches[kTimedKernels] = {};
};

namespace {

// Waits for a stream on every return path: an upload that fails after its
// copies were enqueued still returns only once they are done, so the caller
// may free its host arrays whatever the call returns.
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

// Device scratch of one call, freed on every return path.
struct DeviceTemp {
    void *p = nullptr;
    DeviceTemp() = default;
    DeviceTemp(const DeviceTemp &) = delete;
    DeviceTemp &operator=(const DeviceTemp &) = delete;
    ~DeviceTemp() { if (p) (void)hipFree(p); }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

int fail(wc_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, e_ == hipErrorOutOfMemory ? WC_ERR_NOMEM : WC_ERR_HIP, "%s: %s (%s:%d)", #call, \
                        hipGetErrorString(e_), __FILE__, __LINE__);                               \
    } while (0)

size_t field_count(const wc_ctx *c, int f) {
    const size_t plane = (size_t)c->eta * c->xi;
    return f == 0 ? plane : (f <= 2 ? plane * c->s_rho : plane * c->s_w);
}

double std_min(double a, double b) { return (b < a) ? b : a; }
double std_max(double a, double b) { return (a < b) ? b : a; }

// NumericalHelpers::restartElapsed (NumericalHelpers.hpp:45-52)
double restart_elapsed(double s, double e, double c) {
    if (e > s) {
        if (e <= c) return NAN;
        return std_min(e - s, std_max(0.0, c - s));
    }
    if (e >= c) return NAN;
    return std_min(s - e, std_max(0.0, s - c));
}

// Per-interval scalars of Particle::move (Particle.cpp:81-125, 507, 518-523).
// Returns false when the interval is wholly before a restart checkpoint (the
// reference returns before touching the particle, Particle.cpp:117-120).
bool make_params(const wc_config_data *cfg, double start, double end, int s_w, int s_rho, int eta, int xi,
                 MoveParams &P) {
    P.dti = cfg->dti;
    P.tau0 = cfg->tau0;
    P.survprob = cfg->survprob;
    P.crid = cfg->crid;
    P.sv = cfg->sv;
    P.sigma = cfg->sigma;
    P.shoreLimit = cfg->shoreLimit;
    P.intervalStart = start;
    P.intervalLength = std::fabs(end - start);
    P.elapsed0 = 0;
    if (!std::isnan(cfg->restartCheckpoint)) {
        double e = restart_elapsed(start, end, cfg->restartCheckpoint);
        if (std::isnan(e)) return false;
        P.elapsed0 = e;
    }
    P.kLowerLimit = -s_w + 2;
    P.direction = cfg->trackingDirection;
    P.ageKill = cfg->survprob > 0 ? -cfg->tau0 * std::log(cfg->survprob) : INFINITY;
    // The FP64 move may defer exp(-age/tau0) while age < ageLazy: there
    // -age/tau0 > -L (1 - 1e-6) with L = min(-log(survprob), 700), so the
    // correctly rounded quotient and a faithful exp give a value above
    // survprob by a relative 1e-6 L and far above 0; neither survival test
    // (health < survprob, health > 0) can then depend on the exact value.
    // Anything else (tau0 <= 0 or not finite, survprob >= 1) never defers.
    {
        double L = cfg->survprob > 0 ? -std::log(cfg->survprob) : 700.0;
        if (L > 700.0) L = 700.0;
        const bool ok = std::isfinite(cfg->tau0) && cfg->tau0 > 0 && L > 0;
        P.ageLazy = ok ? cfg->tau0 * L * (1.0 - 1e-6) : -INFINITY;
    }
    P.invDti = 1.0 / cfg->dti;
    P.seed = cfg->randomSeed;
    P.intervalKey = (int64_t)std::llround(std_min(start, end));
    P.jibeHourly = cfg->leewayJibeProbabilityHourly;
    P.jibe = P.jibeHourly > 0;
    P.upper = cfg->upperClosure;
    P.lower = cfg->lowerClosure;
    P.horizontal = cfg->horizontalClosure;
    P.dirInt = cfg->trackingDirection;
    P.randomOn = cfg->random && (P.direction > 0 || cfg->backwardDiffusion);
    P.midpoint = cfg->integrator == WC_INTEGRATOR_MIDPOINT;
    P.s_w = s_w;
    P.s_rho = s_rho;
    P.eta = eta;
    P.xi = xi;
    return true;
}

// The hourly side flip of non-passive particles (Particle.cpp:805-813) runs
// on both move paths; the ensemble's particles are passive, so it has nothing
// to act on there.
int check_supported(wc_ctx *c, const wc_config_data *cfg) {
    if (cfg->driftModel == 1)
        return fail(c, WC_ERR_UNSUPPORTED, "driftModel==1 (leeway) is not implemented on the device path");
    if (!(cfg->dti > 0)) return fail(c, WC_ERR_INVALID, "dti must be > 0");
    if (cfg->trackingDirection != 1 && cfg->trackingDirection != -1)
        return fail(c, WC_ERR_INVALID, "trackingDirection must be +1 or -1");
    return WC_OK;
}

void free_soa(SoA &s) {
    hipFree(s.id); hipFree(s.k); hipFree(s.j); hipFree(s.i); hipFree(s.health); hipFree(s.age); hipFree(s.time);
    hipFree(s.tag); hipFree(s.ev);
    s = SoA{};
}

int alloc_soa(wc_ctx *c, SoA &s, size_t cap) {
    HIPCHK(c, hipMalloc(&s.id, cap * sizeof(uint64_t)));
    HIPCHK(c, hipMalloc(&s.k, cap * sizeof(double)));
    HIPCHK(c, hipMalloc(&s.j, cap * sizeof(double)));
    HIPCHK(c, hipMalloc(&s.i, cap * sizeof(double)));
    HIPCHK(c, hipMalloc(&s.health, cap * sizeof(double)));
    HIPCHK(c, hipMalloc(&s.age, cap * sizeof(double)));
    HIPCHK(c, hipMalloc(&s.time, cap * sizeof(double)));
    HIPCHK(c, hipMalloc(&s.tag, cap * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&s.ev, cap));
    return WC_OK;
}

int reserve(wc_ctx *c, size_t need) {
    if (need <= c->cap) return WC_OK;
    if (need > (size_t)kMaxParticles)
        return fail(c, WC_ERR_INVALID, "at most %d particles per context", kMaxParticles);
    size_t cap = c->cap ? c->cap : 1024;
    while (cap < need) cap *= 2;
    if (cap > (size_t)kMaxParticles) cap = kMaxParticles;
    // every new buffer is allocated before anything is committed, so a
    // failure leaves the context as it was
    SoA a{}, b{};
    uint32_t *blockCounts = nullptr, *perm = nullptr, *perm2 = nullptr, *remap = nullptr;
    auto drop = [&]() {
        free_soa(a);
        free_soa(b);
        hipFree(blockCounts); hipFree(perm); hipFree(perm2); hipFree(remap);
    };
    int rc = alloc_soa(c, a, cap);
    if (!rc) rc = alloc_soa(c, b, cap);
    if (!rc && (hipMalloc(&blockCounts, ((cap + 255) / 256) * sizeof(uint32_t)) != hipSuccess ||
                hipMalloc(&perm, cap * sizeof(uint32_t)) != hipSuccess ||
                hipMalloc(&perm2, cap * sizeof(uint32_t)) != hipSuccess ||
                hipMalloc(&remap, cap * sizeof(uint32_t)) != hipSuccess))
        rc = fail(c, WC_ERR_NOMEM, "particle buffers for %zu particles", cap);
    if (rc) {
        drop();
        return rc;
    }
    if (c->n) {
        const size_t n = c->n;
        const double *src[7] = {c->s.k, c->s.j, c->s.i, c->s.health, c->s.age, c->s.time, nullptr};
        double *dst[7] = {a.k, a.j, a.i, a.health, a.age, a.time, nullptr};
        hipError_t e = hipMemcpyAsync(a.id, c->s.id, n * 8, hipMemcpyDeviceToDevice, c->stream);
        for (int q = 0; q < 6 && e == hipSuccess; ++q)
            e = hipMemcpyAsync(dst[q], src[q], n * 8, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(a.tag, c->s.tag, n * 4, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(a.ev, c->s.ev, n, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            hipStreamSynchronize(c->stream);
            drop();
            return fail(c, WC_ERR_HIP, "copying the population into larger buffers: %s", hipGetErrorString(e));
        }
    }
    free_soa(c->s);
    free_soa(c->s2);
    hipFree(c->d_blockCounts);
    hipFree(c->d_perm); hipFree(c->d_perm2); hipFree(c->d_remap);
    c->s = a;
    c->s2 = b;
    c->cap = cap;
    c->d_blockCounts = blockCounts;
    c->d_perm = perm;
    c->d_perm2 = perm2;
    c->d_remap = remap;
    c->permValid = false;
    return WC_OK;
}

// Host AoS -> device SoA at offset `at`.  Non-passive particles (a drift
// object type other than PASSIVE) move passively with driftModel 0, as in the
// reference; their type and side travel with them.
int upload_particles(wc_ctx *c, const wc_particle_data *p, size_t n, size_t at) {
    std::vector<uint64_t> id(n);
    std::vector<double> f[6];
    std::vector<uint32_t> tag(n);
    for (auto &v : f) v.resize(n);
    for (size_t q = 0; q < n; ++q) {
        id[q] = p[q].id;
        f[0][q] = p[q].k; f[1][q] = p[q].j; f[2][q] = p[q].i;
        f[3][q] = p[q].health; f[4][q] = p[q].age; f[5][q] = p[q].time;
        tag[q] = make_tag(p[q].driftObjectType, p[q].driftSide);
    }
    double *dst[6] = {c->s.k, c->s.j, c->s.i, c->s.health, c->s.age, c->s.time};
    HIPCHK(c, hipMemcpyAsync(c->s.id + at, id.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    for (int q = 0; q < 6; ++q)
        HIPCHK(c, hipMemcpyAsync(dst[q] + at, f[q].data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->s.tag + at, tag.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->s.ev + at, 0, n, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

hipEvent_t take_event(wc_ctx *c) {
    if (!c->spare.empty()) {
        hipEvent_t e = c->spare.back();
        c->spare.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Drops the start event of a timed launch that failed.
void timing_abandon(wc_ctx *c) {
    if (c->cur) c->spare.push_back(c->cur);
    c->cur = nullptr;
}

// Event pair around the next launches on stream s (the compute stream by
// default); a start event left by a failed launch is recycled first.
void timing_begin(wc_ctx *c, hipStream_t s = nullptr) {
    timing_abandon(c);
    if (!c->timing) return;
    c->cur = take_event(c);
    if (c->cur && hipEventRecord(c->cur, s ? s : c->stream) != hipSuccess) {
        c->spare.push_back(c->cur);
        c->cur = nullptr;
    }
}

int timing_end(wc_ctx *c, int kernel, hipStream_t s = nullptr) {
    if (!c->timing || !c->cur) return WC_OK;
    hipEvent_t e = take_event(c);
    if (!e) {
        timing_abandon(c);
        return fail(c, WC_ERR_HIP, "hipEventCreate failed");
    }
    if (hipEventRecord(e, s ? s : c->stream) != hipSuccess) {
        c->spare.push_back(e);
        timing_abandon(c);
        return fail(c, WC_ERR_HIP, "hipEventRecord failed");
    }
    c->pending[kernel].emplace_back(c->cur, e);
    c->cur = nullptr;
    return WC_OK;
}

// Resolves recorded event pairs into the accumulated totals.
int timing_collect(wc_ctx *c) {
    for (int k = 0; k < kTimedKernels; ++k) {
        for (auto &pr : c->pending[k]) {
            HIPCHK(c, hipEventSynchronize(pr.second));
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, pr.first, pr.second));
            c->ms[k] += ms;
            c->launches[k] += 1;
            c->spare.push_back(pr.first);
            c->spare.push_back(pr.second);
        }
        c->pending[k].clear();
    }
    return WC_OK;
}

// Brings the cell-tile order up to date before a move (wc_set_sort): a new
// sort every sortPeriod-th move, or when the population changed in a way
// the order does not follow (set, a new grid).  Timed as kernel group 3.
int sort_population(wc_ctx *c) {
    if (c->sortTile <= 0 || c->n == 0) return WC_OK;
    if (!c->permValid || c->sortCount % c->sortPeriod == 0) {
        timing_begin(c);
        HIPCHK(c, launch_sort_perm(c->s, (int)c->n, c->eta, c->xi, c->sortTile, c->permValid ? c->d_perm : nullptr,
                                   c->d_perm2, c->d_sortBins, c->d_total, c->d_scanSums, c->stream));
        std::swap(c->d_perm, c->d_perm2);
        c->permValid = true;
        const int rc = timing_end(c, 3);
        if (rc) return rc;
    }
    ++c->sortCount;
    return WC_OK;
}

// The {dx, dy} table of the drop-in entry point's lon/lat grid, one per
// device: built on the host with the expression of Particle.cpp:544-574 (the
// table the context path uses, so the two paths give the same bits) the
// first time a grid is seen, and reused while the grid's content hash, taken
// on the device at every call, stays the same.
struct DropinMetric {
    unsigned long long *d_hash = nullptr;
    double2 *table = nullptr;
    size_t cells = 0;
    int eta = 0, xi = 0;
    unsigned long long hash = 0;
    bool valid = false;
    // the interval pack of the drop-in move, rewritten by every call: a call
    // on any stream waits (on the device) for the previous call's move
    void *pack = nullptr;
    size_t packBytes = 0;
    hipEvent_t packFree = nullptr;
};
constexpr int kDropinDevices = 64;
std::mutex dropinLock;
DropinMetric dropinMetric[kDropinDevices];

// The current device's drop-in pack buffer, at least parity_pack_bytes(s_w,
// plane) bytes, after the previous user's move on any stream (the caller
// records packFree after its own launches).  Caller holds dropinLock.
int dropin_pack(int s_w, size_t plane, hipStream_t st, void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return WC_ERR_HIP;
    if (dev < 0 || dev >= kDropinDevices) return WC_ERR_INVALID;
    DropinMetric &M = dropinMetric[dev];
    if (!M.packFree && hipEventCreateWithFlags(&M.packFree, hipEventDisableTiming) != hipSuccess) return WC_ERR_HIP;
    const size_t need = parity_pack_bytes(s_w, plane);
    if (M.packBytes < need) {
        if (hipEventSynchronize(M.packFree) != hipSuccess) return WC_ERR_HIP;   // no launch reads the old buffer
        hipFree(M.pack);
        M.pack = nullptr;
        M.packBytes = 0;
        if (hipMalloc(&M.pack, need) != hipSuccess) return WC_ERR_NOMEM;
        M.packBytes = need;
    }
    if (hipStreamWaitEvent(st, M.packFree, 0) != hipSuccess) return WC_ERR_HIP;
    *out = M.pack;
    return WC_OK;
}

// Caller holds dropinLock.
int dropin_metric(int eta, int xi, const double *lon, const double *lat, hipStream_t st, const double2 **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return WC_ERR_HIP;
    if (dev < 0 || dev >= kDropinDevices) return WC_ERR_INVALID;
    DropinMetric &M = dropinMetric[dev];
    const size_t plane = (size_t)eta * xi;
    if (!M.d_hash && hipMalloc(&M.d_hash, sizeof(unsigned long long)) != hipSuccess) return WC_ERR_NOMEM;
    unsigned long long h = 0;
    if (launch_fill_u32(reinterpret_cast<uint32_t *>(M.d_hash), 2, 0u, st) != hipSuccess ||
        launch_grid_hash(lon, lat, plane, M.d_hash, st) != hipSuccess ||
        hipMemcpyAsync(&h, M.d_hash, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return WC_ERR_HIP;
    if (!(M.valid && M.hash == h && M.eta == eta && M.xi == xi)) {
        std::vector<double> hlon(plane), hlat(plane), dx(plane), dy(plane);
        if (hipMemcpy(hlon.data(), lon, plane * 8, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hlat.data(), lat, plane * 8, hipMemcpyDeviceToHost) != hipSuccess)
            return WC_ERR_HIP;
        wc_cell_metric_table(eta, xi, hlon.data(), hlat.data(), dx.data(), dy.data());
        std::vector<double2> tab(plane);
        for (size_t q = 0; q < plane; ++q) tab[q] = make_double2(dx[q], dy[q]);
        M.valid = false;
        // earlier launches may still read the old table
        if (hipDeviceSynchronize() != hipSuccess) return WC_ERR_HIP;
        if (M.cells != plane) {
            hipFree(M.table);
            M.table = nullptr;
            M.cells = 0;
            if (hipMalloc(&M.table, plane * sizeof(double2)) != hipSuccess) return WC_ERR_NOMEM;
            M.cells = plane;
        }
        if (hipMemcpy(M.table, tab.data(), plane * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
            return WC_ERR_HIP;
        M.hash = h;
        M.eta = eta;
        M.xi = xi;
        M.valid = true;
    }
    *out = M.table;
    return WC_OK;
}

}  // namespace

extern "C" {

const char *wc_version(void) { return "wacomm-hip 0.1 (gfx950)"; }

// Haversine cell sizes, the same expression as Particle.cpp:548-574, for
// every cell (j < eta-1, i < xi-1); edge rows/columns hold 0.
void wc_cell_metric_table(int eta, int xi, const double *lonRad, const double *latRad, double *dx, double *dy) {
    for (int j = 0; j < eta; ++j)
        for (int i = 0; i < xi; ++i) {
            const size_t c = (size_t)j * xi + i;
            dx[c] = 0;
            dy[c] = 0;
            if (j + 1 >= eta || i + 1 >= xi) continue;
            double dLat = latRad[c + xi] - latRad[c];
            double dLon = 0;
            double t = pow(sin(0.5 * dLat), 2) + pow(sin(0.5 * dLon), 2) * cos(latRad[c + xi]) * cos(latRad[c]);
            dy[c] = 2.0 * atan2(pow(t, .5), pow(1.0 - t, .5)) * 6371000.0;
            dLat = 0;
            dLon = lonRad[c + 1] - lonRad[c];
            t = pow(sin(0.5 * dLat), 2) + pow(sin(0.5 * dLon), 2) * cos(latRad[c + 1]) * cos(latRad[c]);
            dx[c] = 2.0 * atan2(pow(t, .5), pow(1.0 - t, .5)) * 6371000.0;
        }
}

void wc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const Philox4 r = philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int q = 0; q < 4; ++q) out[q] = r.v[q];
}

// ---- drop-in entry points --------------------------------------------------

int wc_hip_move_particles(wc_config_data *config, wc_particle_data *particles, int timeIndex, int oceanTime, int sW,
                          int sRho, int eta, int xi, double *times, double *mask, double *lonRad, double *latRad,
                          double *depthIntervals, double *h, float *zeta, float *u, float *v, float *w, float *akt,
                          float *windU10, float *windV10, float *stokesU, float *stokesV, int particleCount,
                          int numThread, int numGPU, void *stream) {
    (void)mask; (void)numThread; (void)numGPU;
    (void)windU10; (void)windV10; (void)stokesU; (void)stokesV;
    if (!config || !times || !lonRad || !latRad || !depthIntervals || !h || !zeta || !u || !v || !w || !akt)
        return WC_ERR_INVALID;
    if (particleCount < 0 || particleCount > kMaxParticles || oceanTime <= 0 || timeIndex < 0 ||
        timeIndex >= oceanTime || sW != sRho + 1 || sW < 2 || eta < 2 || xi < 2)
        return WC_ERR_INVALID;
    if (particleCount == 0) return WC_OK;
    if (!particles) return WC_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    // config and times live in device memory, as in the reference call
    wc_config_data cfg;
    if (hipMemcpyAsync(&cfg, config, sizeof cfg, hipMemcpyDeviceToHost, st) != hipSuccess) return WC_ERR_HIP;
    std::vector<double> t(oceanTime);
    if (hipMemcpyAsync(t.data(), times, oceanTime * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
        return WC_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return WC_ERR_HIP;
    int rc = check_supported(nullptr, &cfg);
    if (rc) return rc;
    // interval and record pair: Particle.cpp:110-115, 183-184
    const int dir = cfg.trackingDirection;
    int next = timeIndex + dir;
    const double start = t[timeIndex];
    const double end = (next >= 0 && next < oceanTime) ? t[next] : start + dir * cfg.deltat;
    if (next < 0 || next >= oceanTime) next = timeIndex;
    MoveParams P;
    if (!make_params(&cfg, start, end, sW, sRho, eta, xi, P)) return WC_OK;
    const size_t plane = (size_t)eta * xi;
    std::lock_guard<std::mutex> guard(dropinLock);
    const double2 *metric = nullptr;
    rc = dropin_metric(eta, xi, lonRad, latRad, st, &metric);
    if (rc) return rc;
    RecordPair<float> F{zeta + plane * timeIndex, zeta + plane * next,
                        u + plane * sRho * timeIndex, u + plane * sRho * next,
                        v + plane * sRho * timeIndex, v + plane * sRho * next,
                        w + plane * sW * timeIndex, w + plane * sW * next,
                        akt + plane * sW * timeIndex, akt + plane * sW * next};
    GridPtrs G{h, depthIntervals, nullptr, nullptr, lonRad, latRad, nullptr, nullptr, nullptr, nullptr};
    void *pack = nullptr;
    rc = dropin_pack(sW, plane, st, &pack);
    if (rc) return rc;
    if (launch_move_parity_aos(P, F, G, metric, pack, particles, particleCount, nullptr, st) != hipSuccess)
        return WC_ERR_HIP;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipEventRecord(dropinMetric[dev].packFree, st) != hipSuccess)
        return WC_ERR_HIP;
    return WC_OK;
}

int wc_hip_concentration(const wc_particle_data *particles, int particleCount, int s_rho, int eta, int xi,
                         float *conc, void *stream) {
    if (particleCount < 0 || particleCount > kMaxParticles || s_rho < 1 || eta < 1 || xi < 1 || !conc)
        return WC_ERR_INVALID;
    if ((size_t)s_rho * eta * xi > 0xFFFFFFFFull) return WC_ERR_INVALID;
    if (particleCount == 0) return WC_OK;
    if (!particles) return WC_ERR_INVALID;
    return launch_grid_count_aos_float(particles, particleCount, s_rho, eta, xi, conc, (hipStream_t)stream) ==
                   hipSuccess
               ? WC_OK
               : WC_ERR_HIP;
}

// ---- native context ----------------------------------------------------------

int wc_ctx_create(int device, wc_ctx **out) {
    if (!out) return WC_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return WC_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return WC_ERR_NO_DEVICE;
    wc_ctx *c = new wc_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipMalloc(&c->d_total, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&c->d_steps, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&c->d_scanSums, kScanSumsCap * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(c->d_steps, 0, sizeof(unsigned long long)) != hipSuccess) {
        wc_ctx_destroy(c);
        return WC_ERR_HIP;
    }
    *out = c;
    return WC_OK;
}

static void free_ens_graph(wc_ctx *c) {
    if (c->ens.exec) hipGraphExecDestroy(c->ens.exec);
    if (c->ens.graph) hipGraphDestroy(c->ens.graph);
    c->ens.exec = nullptr;
    c->ens.graph = nullptr;
}

static void free_ens_soa(EnsSoA &s) {
    hipFree(s.id); hipFree(s.cell); hipFree(s.fj); hipFree(s.fi); hipFree(s.k); hipFree(s.health);
    hipFree(s.age); hipFree(s.time); hipFree(s.scen);
    s = EnsSoA{};
}

static void free_ens(wc_ctx *c) {
    free_ens_graph(c);
    free_ens_soa(c->ens.s);
    free_ens_soa(c->ens.save);
    free_ens_soa(c->ens.tmp);
    hipFree(c->ens.d_sigma); hipFree(c->ens.d_surface); hipFree(c->ens.d_steps); hipFree(c->ens.d_stepIdx);
    hipFree(c->ens.d_count); hipFree(c->ens.d_bins);
    const int block = c->ens.block, ppt = c->ens.ppt, occ = c->ens.occ;
    const int tile = c->ens.sortTile, period = c->ens.sortPeriod;
    const bool tuned = c->ens.tuned;
    c->ens = wc_ctx::Ens{};
    c->ens.block = block; c->ens.ppt = ppt; c->ens.occ = occ; c->ens.tuned = tuned;
    c->ens.sortTile = tile; c->ens.sortPeriod = period;
}

static void free_grid(wc_ctx *c) {
    free_ens_graph(c);      // the captured step holds the pack buffer
    hipFree(c->d_h); hipFree(c->d_depth); hipFree(c->d_dx); hipFree(c->d_dy); hipFree(c->d_mask);
    c->d_h = c->d_depth = c->d_dx = c->d_dy = c->d_mask = nullptr;
    hipFree(c->d_hf); hipFree(c->d_md); hipFree(c->d_depthf); hipFree(c->d_pack); hipFree(c->d_metric);
    c->d_metric = nullptr;
    c->d_hf = c->d_md = c->d_depthf = nullptr;
    c->d_pack = nullptr;
    for (auto &s : c->slots) {
        for (int f = 0; f < 5; ++f) hipFree(s.f[f]);
        if (s.released) hipEventDestroy(s.released);
    }
    c->slots.clear();
    hipFree(c->d_counts);
    hipFree(c->d_conc);
    hipFree(c->d_face); hipFree(c->d_rot);
    hipFree(c->d_u_raw); hipFree(c->d_v_raw); hipFree(c->d_flag);
    c->d_face = nullptr;
    c->d_rot = nullptr;
    c->d_u_raw = c->d_v_raw = nullptr;
    c->d_flag = nullptr;
    hipFree(c->d_key);
    c->d_key = nullptr;
    c->haveRoms = false;
    hipFree(c->d_lat); hipFree(c->d_lon); hipFree(c->d_sw);
    c->d_lat = c->d_lon = c->d_sw = nullptr;
    c->haveGeo = false;
    c->d_counts = nullptr;
    c->d_conc = nullptr;
    c->haveGrid = false;
}

int wc_ctx_destroy(wc_ctx *c) {
    if (!c) return WC_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->copy) hipStreamSynchronize(c->copy);
    free_grid(c);
    free_ens(c);
    free_soa(c->s);
    free_soa(c->s2);
    hipFree(c->d_blockCounts);
    hipFree(c->d_perm); hipFree(c->d_perm2); hipFree(c->d_remap);
    hipFree(c->d_scanSums);
    hipFree(c->d_total);
    hipFree(c->d_steps);
    hipFree(c->d_rec);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    for (int k = 0; k < kTimedKernels; ++k)
        for (auto &pr : c->pending[k]) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
    for (auto e : c->spare) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->copy) hipStreamDestroy(c->copy);
    delete c;
    return WC_OK;
}

const char *wc_last_error(const wc_ctx *c) { return c ? c->err.c_str() : ""; }

int wc_sync(wc_ctx *c) {
    if (!c) return WC_ERR_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

int wc_set_grid(wc_ctx *c, int s_w, int s_rho, int eta, int xi, const double *lonRad, const double *latRad,
                const double *h, const double *mask, const double *depthIntervals, int n_slots, int store) {
    if (!c) return WC_ERR_INVALID;
    if (s_w != s_rho + 1 || s_rho < 1 || eta < 2 || xi < 2 || n_slots < 2 || n_slots > 4096 || !lonRad ||
        !latRad || !h || !depthIntervals)
        return fail(c, WC_ERR_INVALID, "invalid grid: s_w=%d s_rho=%d eta=%d xi=%d slots=%d", s_w, s_rho, eta, xi,
                    n_slots);
    if (store < WC_STORE_F32 || store > WC_STORE_I16_TILE)
        return fail(c, WC_ERR_INVALID, "unknown storage type %d", store);
    if ((size_t)s_w * eta * xi > 0xFFFFFFFFull) return fail(c, WC_ERR_INVALID, "grid too large");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_grid(c);
    c->s_w = s_w; c->s_rho = s_rho; c->eta = eta; c->xi = xi; c->store = store;
    const size_t plane = (size_t)eta * xi;
    std::vector<double> dx(plane), dy(plane), m(plane, 1.0);
    wc_cell_metric_table(eta, xi, lonRad, latRad, dx.data(), dy.data());
    if (mask) std::memcpy(m.data(), mask, plane * sizeof(double));
    HIPCHK(c, hipMalloc(&c->d_h, plane * 8));
    HIPCHK(c, hipMalloc(&c->d_dx, plane * 8));
    HIPCHK(c, hipMalloc(&c->d_dy, plane * 8));
    HIPCHK(c, hipMalloc(&c->d_mask, plane * 8));
    HIPCHK(c, hipMalloc(&c->d_depth, (size_t)(s_w - 1) * 8));
    HIPCHK(c, hipMemcpy(c->d_h, h, plane * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_dx, dx.data(), plane * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_dy, dy.data(), plane * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_mask, m.data(), plane * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_depth, depthIntervals, (size_t)(s_w - 1) * 8, hipMemcpyHostToDevice));
    // FP32 tables of the mixed-precision path
    std::vector<float> hf(plane), md(2 * plane), depthf(s_w - 1);
    for (size_t q = 0; q < plane; ++q) {
        hf[q] = (float)h[q];
        md[2 * q] = dx[q] > 0 ? (float)(1.0 / dx[q]) : 0.0f;
        md[2 * q + 1] = dy[q] > 0 ? (float)(1.0 / dy[q]) : 0.0f;
    }
    for (int q = 0; q < s_w - 1; ++q) depthf[q] = (float)depthIntervals[q];
    HIPCHK(c, hipMalloc(&c->d_hf, plane * 4));
    HIPCHK(c, hipMalloc(&c->d_md, plane * 8));
    HIPCHK(c, hipMalloc(&c->d_depthf, (size_t)(s_w - 1) * 4));
    {
        const size_t a = fast_pack_bytes(store, s_w, eta, xi), b = parity_pack_bytes(s_w, plane);
        HIPCHK(c, hipMalloc(&c->d_pack, a > b ? a : b));
    }
    std::vector<double> metric(2 * plane);
    for (size_t q = 0; q < plane; ++q) { metric[2 * q] = dx[q]; metric[2 * q + 1] = dy[q]; }
    HIPCHK(c, hipMalloc(&c->d_metric, plane * 16));
    HIPCHK(c, hipMemcpy(c->d_metric, metric.data(), plane * 16, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_hf, hf.data(), plane * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_md, md.data(), plane * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_depthf, depthf.data(), (size_t)(s_w - 1) * 4, hipMemcpyHostToDevice));
    c->slots.resize(n_slots);
    for (auto &s : c->slots) {
        for (int f = 0; f < 5; ++f) HIPCHK(c, hipMalloc(&s.f[f], field_count(c, f) * sizeof(float)));
        HIPCHK(c, hipEventCreateWithFlags(&s.released, hipEventDisableTiming));
        HIPCHK(c, hipEventRecord(s.released, c->stream));
    }
    HIPCHK(c, hipMalloc(&c->d_counts, plane * s_rho * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&c->d_conc, plane * s_rho * sizeof(float)));
    HIPCHK(c, hipMalloc(&c->d_sortBins, (plane + 1) * sizeof(uint32_t)));     // tiles of one cell, plus one
    c->permValid = false;
    c->haveGrid = true;
    return WC_OK;
}

int wc_upload_record(wc_ctx *c, int slot, const float *zeta, const float *u, const float *v, const float *w,
                     const float *akt) {
    if (!c) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    if (slot < 0 || slot >= (int)c->slots.size()) return fail(c, WC_ERR_INVALID, "slot %d out of range", slot);
    const float *src[5] = {zeta, u, v, w, akt};
    for (int f = 0; f < 5; ++f)
        if (!src[f]) return fail(c, WC_ERR_INVALID, "null forcing field %d", f);
    HIPCHK(c, hipSetDevice(c->device));
    auto &s = c->slots[slot];
    // the copy waits only for the enqueued work that uses this slot; work on
    // other slots (the move of the current interval) runs meanwhile.  The
    // call returns when the copy is done, so the caller may reuse its arrays,
    // and every later launch sees the record.
    s.filled = false;
    HIPCHK(c, hipStreamWaitEvent(c->copy, s.released, 0));
    StreamDrain drain{c->copy};
    for (int f = 0; f < 5; ++f)
        HIPCHK(c, hipMemcpyAsync(s.f[f], src[f], field_count(c, f) * sizeof(float), hipMemcpyHostToDevice, c->copy));
    HIPCHK(c, hipStreamSynchronize(c->copy));
    s.filled = true;
    s.record = -1;
    return WC_OK;
}

// validRomsValue (ROMSAdapter.cpp:45-49) over the attribute list plus the two
// NetCDF fill values romsMissingValues appends (:40-41)
static bool roms_valid(double value, const double *extra, int n_extra) {
    if (!std::isfinite(value)) return false;
    for (int q = 0; q < n_extra; ++q)
        if (value == extra[q]) return false;
    return value != (double)kNcFillFloat && value != kNcFillDouble;
}

int wc_set_roms_grid(wc_ctx *c, const wc_roms_grid *g, int n_slots, int store) {
    if (!c || !g) return WC_ERR_INVALID;
    const int s_rho = g->s_rho, eta = g->eta_rho, xi = g->xi_rho;
    if (s_rho < 1 || eta < 2 || xi < 2 || g->eta_u < 0 || g->xi_u < 0 || g->eta_v < 0 || g->xi_v < 0)
        return fail(c, WC_ERR_INVALID, "invalid ROMS dimensions");
    if (g->eta_u > eta || g->xi_u > xi || g->eta_v > eta || g->xi_v > xi)
        return fail(c, WC_ERR_INVALID, "u/v grids (%dx%d, %dx%d) larger than the rho grid (%dx%d)", g->eta_u,
                    g->xi_u, g->eta_v, g->xi_v, eta, xi);
    if (!g->s_w || !g->lon_rho || !g->lat_rho || !g->h || !g->mask_rho || (g->eta_u * g->xi_u && !g->mask_u) ||
        (g->eta_v * g->xi_v && !g->mask_v) || (!g->earth_relative && !g->angle))
        return fail(c, WC_ERR_INVALID, "null ROMS grid field");
    if (g->n_missing_u < 0 || g->n_missing_v < 0 || g->n_missing_angle < 0 ||
        g->n_missing_u > WC_ROMS_MAX_MISSING || g->n_missing_v > WC_ROMS_MAX_MISSING ||
        g->n_missing_angle > WC_ROMS_MAX_MISSING ||
        (g->n_missing_u && !g->missing_u) || (g->n_missing_v && !g->missing_v) ||
        (g->n_missing_angle && !g->missing_angle))
        return fail(c, WC_ERR_INVALID, "at most %d missing values per variable", WC_ROMS_MAX_MISSING);
    const size_t plane = (size_t)eta * xi, pu = (size_t)g->eta_u * g->xi_u, pv = (size_t)g->eta_v * g->xi_v;
    // input checks of ROMSAdapter.cpp:146-166 that do not depend on a record
    for (size_t q = 0; q < plane; ++q) {
        const double m = g->mask_rho[q];
        if (m != 0.0 && m != 1.0) return fail(c, WC_ERR_INVALID, "ROMS rho mask requires finite zero/one values");
        if (!g->earth_relative && m > 0.0 &&
            !roms_valid(g->angle[q], g->missing_angle, g->n_missing_angle))
            return fail(c, WC_ERR_INVALID, "ROMS angle is missing or non-finite at a wet rho point (%zu, %zu)",
                        q / xi, q % xi);
    }
    for (size_t q = 0; q < pu; ++q)
        if (g->mask_u[q] != 0.0 && g->mask_u[q] != 1.0)
            return fail(c, WC_ERR_INVALID, "ROMS u mask requires finite zero/one values");
    for (size_t q = 0; q < pv; ++q)
        if (g->mask_v[q] != 0.0 && g->mask_v[q] != 1.0)
            return fail(c, WC_ERR_INVALID, "ROMS v mask requires finite zero/one values");
    // static fields of ROMSAdapter.cpp:232-275
    const int s_w = s_rho + 1;
    std::vector<double> lonRad(plane), latRad(plane), h(plane), depth(s_w - 1);
    for (int k = -s_w + 2; k <= 0; ++k) depth[k + s_w - 2] = g->s_w[k + s_w - 1] - g->s_w[k + s_w - 2];
    for (size_t q = 0; q < plane; ++q) {
        h[q] = g->mask_rho[q] == 0 ? 0.0 : g->h[q];
        lonRad[q] = 0.0174533 * g->lon_rho[q];
        latRad[q] = 0.0174533 * g->lat_rho[q];
    }
    int rc = wc_set_grid(c, s_w, s_rho, eta, xi, lonRad.data(), latRad.data(), h.data(), g->mask_rho, depth.data(),
                         n_slots, store);
    if (rc) return rc;
    // device tables: wet/face bits per rho point, {cos, sin} of the angle (1, 0 on land)
    const int eta_u = g->eta_u, xi_u = g->xi_u, eta_v = g->eta_v, xi_v = g->xi_v;
    std::vector<uint8_t> face(plane);
    std::vector<double2> rot(plane);
    for (int j = 0; j < eta; ++j)
        for (int i = 0; i < xi; ++i) {
            const size_t q = (size_t)j * xi + i;
            const bool wet = g->mask_rho[q] > 0.0;
            uint32_t f = wet ? kFaceWet : 0;
            if (j < eta_u && i < xi_u && g->mask_u[(size_t)j * xi_u + i] > 0.0) f |= kFaceU;
            if (j < eta_u && i > 0 && i - 1 < xi_u && g->mask_u[(size_t)j * xi_u + i - 1] > 0.0) f |= kFaceUWest;
            if (j < eta_v && i < xi_v && g->mask_v[(size_t)j * xi_v + i] > 0.0) f |= kFaceV;
            if (j > 0 && j - 1 < eta_v && i < xi_v && g->mask_v[(size_t)(j - 1) * xi_v + i] > 0.0) f |= kFaceVSouth;
            face[q] = (uint8_t)f;
            const bool turn = wet && !g->earth_relative;
            rot[q] = make_double2(turn ? std::cos(g->angle[q]) : 1.0, turn ? std::sin(g->angle[q]) : 0.0);
        }
    HIPCHK(c, hipMalloc(&c->d_face, plane));
    HIPCHK(c, hipMalloc(&c->d_rot, plane * sizeof(double2)));
    HIPCHK(c, hipMalloc(&c->d_u_raw, (pu ? pu : 1) * s_rho * sizeof(float)));
    HIPCHK(c, hipMalloc(&c->d_v_raw, (pv ? pv : 1) * s_rho * sizeof(float)));
    HIPCHK(c, hipMalloc(&c->d_flag, sizeof(uint32_t)));
    HIPCHK(c, hipMemcpy(c->d_face, face.data(), plane, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_rot, rot.data(), plane * sizeof(double2), hipMemcpyHostToDevice));
    c->roms = RomsDims{s_rho, eta, xi, g->eta_u, g->xi_u, g->eta_v, g->xi_v, g->earth_relative ? 1 : 0};
    c->romsLon.assign(g->lon_rho, g->lon_rho + plane);
    c->romsLat.assign(g->lat_rho, g->lat_rho + plane);
    c->romsH = h;
    c->romsMask.assign(g->mask_rho, g->mask_rho + plane);
    c->romsSw.assign(g->s_w, g->s_w + s_w);
    if (g->s_rho_values)
        c->romsSrho.assign(g->s_rho_values, g->s_rho_values + s_rho);
    else
        c->romsSrho.clear();
    RomsMissing &m = c->romsMissing;
    m = RomsMissing{};
    for (int q = 0; q < g->n_missing_u; ++q) m.u[m.nu++] = g->missing_u[q];
    for (int q = 0; q < g->n_missing_v; ++q) m.v[m.nv++] = g->missing_v[q];
    m.u[m.nu++] = (double)kNcFillFloat; m.u[m.nu++] = kNcFillDouble;
    m.v[m.nv++] = (double)kNcFillFloat; m.v[m.nv++] = kNcFillDouble;
    c->haveRoms = true;
    return wc_set_geo(c, g->lat_rho, g->lon_rho, g->s_w);
}

int wc_set_geo(wc_ctx *c, const double *lat, const double *lon, const double *s_w) {
    if (!c || !lat || !lon || !s_w) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t plane = (size_t)c->eta * c->xi;
    if (!c->d_lat) HIPCHK(c, hipMalloc(&c->d_lat, plane * sizeof(double)));
    if (!c->d_lon) HIPCHK(c, hipMalloc(&c->d_lon, plane * sizeof(double)));
    if (!c->d_sw) HIPCHK(c, hipMalloc(&c->d_sw, c->s_w * sizeof(double)));
    HIPCHK(c, hipMemcpy(c->d_lat, lat, plane * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_lon, lon, plane * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_sw, s_w, c->s_w * sizeof(double), hipMemcpyHostToDevice));
    c->haveGeo = true;
    return WC_OK;
}

int wc_particle_records(wc_ctx *c, const wc_records *out, size_t capacity, size_t *count) {
    if (!c || !count) return WC_ERR_INVALID;
    *count = 0;
    if (!c->haveGeo) return fail(c, WC_ERR_INVALID, "wc_set_geo or wc_set_roms_grid first");
    if (c->n == 0) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->recCap < c->cap) {
        hipFree(c->d_rec);
        c->d_rec = nullptr;
        c->recCap = 0;
        HIPCHK(c, hipMalloc(&c->d_rec, 3 * c->cap * sizeof(double)));
        c->recCap = c->cap;
    }
    double *dep = c->d_rec, *lat = dep + c->recCap, *lon = lat + c->recCap;
    const GeoPtrs g{c->d_lat, c->d_lon, c->d_h, c->d_sw, c->eta, c->xi, c->s_w};
    HIPCHK(c, launch_particle_records(c->s, (int)c->n, g, c->s2, dep, lat, lon, c->d_blockCounts, c->d_total,
                                      c->stream));
    unsigned long long total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, c->d_total, sizeof total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *count = (size_t)total;
    if (!out || !total) return WC_OK;
    if (capacity < total) return fail(c, WC_ERR_INVALID, "%llu records, capacity %zu", total, capacity);
    const size_t b = total * sizeof(double);
    const struct { void *dst; const void *src; } f[10] = {
        {out->id, c->s2.id}, {out->lat, lat}, {out->lon, lon}, {out->depth, dep}, {out->k, c->s2.k},
        {out->j, c->s2.j}, {out->i, c->s2.i}, {out->health, c->s2.health}, {out->age, c->s2.age},
        {out->time, c->s2.time}};
    for (const auto &e : f)
        if (e.dst) HIPCHK(c, hipMemcpyAsync(e.dst, e.src, b, hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> tag;
    if (out->object_type || out->drift_side) {
        tag.resize(total);
        HIPCHK(c, hipMemcpyAsync(tag.data(), c->s2.tag, total * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < tag.size(); ++q) {
        if (out->object_type) out->object_type[q] = tag_type(tag[q]);
        if (out->drift_side) out->drift_side[q] = tag_side(tag[q]);
    }
    return WC_OK;
}

int wc_upload_roms_record(wc_ctx *c, int slot, const float *zeta, const float *u, const float *v, const float *w,
                          const float *akt) {
    if (!c) return WC_ERR_INVALID;
    if (!c->haveRoms) return fail(c, WC_ERR_INVALID, "wc_set_roms_grid first");
    if (slot < 0 || slot >= (int)c->slots.size()) return fail(c, WC_ERR_INVALID, "slot %d out of range", slot);
    const RomsDims &d = c->roms;
    const size_t pu = (size_t)d.eta_u * d.xi_u * d.s_rho, pv = (size_t)d.eta_v * d.xi_v * d.s_rho;
    if (!zeta || !w || !akt || (pu && !u) || (pv && !v)) return fail(c, WC_ERR_INVALID, "null ROMS record field");
    HIPCHK(c, hipSetDevice(c->device));
    auto &s = c->slots[slot];
    s.filled = false;
    s.record = -1;
    // on the copy stream, after the enqueued work that uses this slot (as
    // wc_upload_record): the copies and the normalisation overlap a move on
    // the other slots
    hipStream_t cs = c->copy;
    HIPCHK(c, hipStreamWaitEvent(cs, s.released, 0));
    StreamDrain drain{cs};
    HIPCHK(c, hipMemcpyAsync(s.f[0], zeta, field_count(c, 0) * sizeof(float), hipMemcpyHostToDevice, cs));
    HIPCHK(c, hipMemcpyAsync(s.f[3], w, field_count(c, 3) * sizeof(float), hipMemcpyHostToDevice, cs));
    HIPCHK(c, hipMemcpyAsync(s.f[4], akt, field_count(c, 4) * sizeof(float), hipMemcpyHostToDevice, cs));
    if (pu) HIPCHK(c, hipMemcpyAsync(c->d_u_raw, u, pu * sizeof(float), hipMemcpyHostToDevice, cs));
    if (pv) HIPCHK(c, hipMemcpyAsync(c->d_v_raw, v, pv * sizeof(float), hipMemcpyHostToDevice, cs));
    HIPCHK(c, launch_fill_u32(c->d_flag, 1, 0u, cs));
    timing_begin(c, cs);
    const hipError_t ne = launch_roms_normalize(d, c->romsMissing, c->d_face, c->d_rot, c->d_u_raw, c->d_v_raw,
                                                s.f[1], s.f[2], s.f[3], s.f[4], s.f[0], c->d_flag, cs);
    if (ne != hipSuccess) {
        timing_abandon(c);
        return fail(c, WC_ERR_HIP, "ROMS normalisation launch: %s", hipGetErrorString(ne));
    }
    int trc = timing_end(c, 2, cs);
    if (trc) return trc;
    uint32_t flag = 0;
    HIPCHK(c, hipMemcpyAsync(&flag, c->d_flag, sizeof flag, hipMemcpyDeviceToHost, cs));
    HIPCHK(c, hipStreamSynchronize(cs));
    if (flag & 1u) return fail(c, WC_ERR_INVALID, "ROMS u is missing or non-finite at a wet face");
    if (flag & 2u) return fail(c, WC_ERR_INVALID, "ROMS v is missing or non-finite at a wet face");
    if (flag & 4u) return fail(c, WC_ERR_INVALID, "ROMS normalized velocity is outside finite float range");
    s.filled = true;
    return WC_OK;
}

int wc_check_roms_grid(wc_ctx *c, const wc_roms_grid *g) {
    if (!c || !g) return WC_ERR_INVALID;
    if (!c->haveRoms) return fail(c, WC_ERR_INVALID, "wc_set_roms_grid first");
    // OceanModelAdapter.cpp:199-214, in the reference's order and with its tolerances
    const int s_rho = c->roms.s_rho, eta = c->roms.eta, xi = c->roms.xi;
    if (g->s_rho != s_rho || g->eta_rho != eta || g->xi_rho != xi)
        return fail(c, WC_ERR_INVALID, "Adjacent forcing files have incompatible grid dimensions");
    if (!g->s_w || !g->lon_rho || !g->lat_rho || !g->h || !g->mask_rho)
        return fail(c, WC_ERR_INVALID, "null ROMS grid field");
    if (g->s_rho_values && !c->romsSrho.empty())
        for (int k = 0; k < s_rho; ++k)
            if (std::fabs(c->romsSrho[k] - g->s_rho_values[k]) > 1e-12)
                return fail(c, WC_ERR_INVALID, "Adjacent forcing files have incompatible vertical coordinates");
    for (int k = 0; k <= s_rho; ++k)
        if (std::fabs(c->romsSw[k] - g->s_w[k]) > 1e-12)
            return fail(c, WC_ERR_INVALID, "Adjacent forcing files have incompatible vertical coordinates");
    const size_t plane = (size_t)eta * xi;
    for (size_t q = 0; q < plane; ++q) {
        const double h = g->mask_rho[q] == 0 ? 0.0 : g->h[q];      // both sides as loaded (:244-252)
        if (std::fabs(c->romsLon[q] - g->lon_rho[q]) > 1e-10 || std::fabs(c->romsLat[q] - g->lat_rho[q]) > 1e-10 ||
            std::fabs(c->romsH[q] - h) > 1e-8 || c->romsMask[q] != g->mask_rho[q])
            return fail(c, WC_ERR_INVALID, "Adjacent forcing files have incompatible horizontal grids");
    }
    return WC_OK;
}

int wc_compare_records(wc_ctx *c, int slot_a, int slot_b, int *conflict, long long *point) {
    if (!c || !conflict) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    const int ns = (int)c->slots.size();
    if (slot_a < 0 || slot_a >= ns || slot_b < 0 || slot_b >= ns || !c->slots[slot_a].filled ||
        !c->slots[slot_b].filled)
        return fail(c, WC_ERR_INVALID, "forcing slots %d/%d not uploaded", slot_a, slot_b);
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->d_key) HIPCHK(c, hipMalloc(&c->d_key, sizeof(unsigned long long)));
    const auto &a = c->slots[slot_a], &b = c->slots[slot_b];
    RecordPair<float> F{a.f[0], b.f[0], a.f[1], b.f[1], a.f[2], b.f[2], a.f[3], b.f[3], a.f[4], b.f[4]};
    HIPCHK(c, launch_fill_u32(reinterpret_cast<uint32_t *>(c->d_key), 2, 0xFFFFFFFFu, c->stream));
    HIPCHK(c, launch_compare_records(F, c->s_rho, (size_t)c->eta * c->xi, c->d_key, c->stream));
    HIPCHK(c, hipEventRecord(a.released, c->stream));
    HIPCHK(c, hipEventRecord(b.released, c->stream));
    unsigned long long key = 0;
    HIPCHK(c, hipMemcpyAsync(&key, c->d_key, sizeof key, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *conflict = key == ~0ull ? 0 : (int)(key % 3) + 1;
    if (point) *point = key == ~0ull ? -1 : (long long)(key / 3);
    return WC_OK;
}

int wc_get_record(wc_ctx *c, int slot, float *zeta, float *u, float *v, float *w, float *akt) {
    if (!c) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    if (slot < 0 || slot >= (int)c->slots.size() || !c->slots[slot].filled)
        return fail(c, WC_ERR_INVALID, "slot %d out of range or not uploaded", slot);
    HIPCHK(c, hipSetDevice(c->device));
    float *dst[5] = {zeta, u, v, w, akt};
    for (int f = 0; f < 5; ++f)
        if (dst[f])
            HIPCHK(c, hipMemcpyAsync(dst[f], c->slots[slot].f[f], field_count(c, f) * sizeof(float),
                                     hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

int wc_set_particles(wc_ctx *c, const wc_particle_data *p, size_t n) {
    if (!c || (n && !p)) return WC_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = reserve(c, n);
    if (rc) return rc;
    c->n = 0;
    c->permValid = false;
    if (n) {
        rc = upload_particles(c, p, n, 0);
        if (rc) return rc;
    }
    c->n = n;
    return WC_OK;
}

int wc_append_particles(wc_ctx *c, const wc_particle_data *p, size_t n) {
    if (!c || (n && !p)) return WC_ERR_INVALID;
    if (!n) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = reserve(c, c->n + n);
    if (rc) return rc;
    rc = upload_particles(c, p, n, c->n);
    if (rc) return rc;
    if (c->permValid) HIPCHK(c, launch_perm_fill(c->d_perm, c->n, c->n + n, c->stream));   // appended last
    c->n += n;
    return WC_OK;
}

int wc_get_particles(wc_ctx *c, wc_particle_data *p, size_t n) {
    if (!c || (n && !p)) return WC_ERR_INVALID;
    if (n > c->n) return fail(c, WC_ERR_INVALID, "asked for %zu particles, context holds %zu", n, c->n);
    if (!n) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<uint64_t> id(n);
    std::vector<double> f[6];
    for (auto &v : f) v.resize(n);
    std::vector<uint32_t> tag(n);
    const double *src[6] = {c->s.k, c->s.j, c->s.i, c->s.health, c->s.age, c->s.time};
    HIPCHK(c, hipMemcpyAsync(id.data(), c->s.id, n * 8, hipMemcpyDeviceToHost, c->stream));
    for (int q = 0; q < 6; ++q)
        HIPCHK(c, hipMemcpyAsync(f[q].data(), src[q], n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(tag.data(), c->s.tag, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < n; ++q) {
        wc_particle_data d{};
        d.id = id[q];
        d.k = f[0][q]; d.j = f[1][q]; d.i = f[2][q];
        d.health = f[3][q]; d.age = f[4][q]; d.time = f[5][q];
        d.driftObjectType = tag_type(tag[q]);
        d.driftSide = tag_side(tag[q]);
        p[q] = d;
    }
    return WC_OK;
}

int wc_particle_count(const wc_ctx *c, size_t *n) {
    if (!c || !n) return WC_ERR_INVALID;
    *n = c->n;
    return WC_OK;
}

int wc_set_event_tracking(wc_ctx *c, int enabled) {
    if (!c) return WC_ERR_INVALID;
    c->trackEvents = enabled != 0;
    return WC_OK;
}

int wc_clear_events(wc_ctx *c) {
    if (!c) return WC_ERR_INVALID;
    if (!c->n) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemsetAsync(c->s.ev, 0, c->n, c->stream));
    return WC_OK;
}

int wc_set_sort(wc_ctx *c, int tile, int period) {
    if (!c) return WC_ERR_INVALID;
    if (tile < 0 || tile > 4096 || period < 1) return fail(c, WC_ERR_INVALID, "sort tile %d / period %d", tile, period);
    c->sortTile = tile;
    c->sortPeriod = period;
    c->sortCount = 0;
    c->permValid = false;
    return WC_OK;
}

int wc_set_lds(wc_ctx *c, int chunk, int halo) {
    if (!c) return WC_ERR_INVALID;
    if (chunk < 0 || chunk > 1024 || halo < 0 || halo > 16)
        return fail(c, WC_ERR_INVALID, "LDS chunk %d (0..1024), halo %d (0..16)", chunk, halo);
    c->lds = LdsTiling{chunk, halo};
    return WC_OK;
}

int wc_get_lds(wc_ctx *c, int *chunk, int *halo) {
    if (!c) return WC_ERR_INVALID;
    if (chunk) *chunk = c->lds.chunk;
    if (halo) *halo = c->lds.halo;
    return WC_OK;
}

int wc_get_sort(wc_ctx *c, int *tile, int *period) {
    if (!c) return WC_ERR_INVALID;
    if (tile) *tile = c->sortTile;
    if (period) *period = c->sortPeriod;
    return WC_OK;
}

int wc_autotune_sort(wc_ctx *c, const wc_config_data *cfg, double start, double end, int slot0, int slot1, int path,
                     int *best_tile, double *best_ms) {
    if (!c || !cfg) return WC_ERR_INVALID;
    if (!c->haveGrid || c->n == 0) return fail(c, WC_ERR_INVALID, "wc_set_grid and particles first");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t n = c->n;
    // the population and the step counter are restored after the trials
    auto copy = [&](const SoA &dst, const SoA &src) -> int {
        HIPCHK(c, hipMemcpyAsync(dst.id, src.id, n * 8, hipMemcpyDeviceToDevice, c->stream));
        double *d[6] = {dst.k, dst.j, dst.i, dst.health, dst.age, dst.time};
        double *s[6] = {src.k, src.j, src.i, src.health, src.age, src.time};
        for (int q = 0; q < 6; ++q) HIPCHK(c, hipMemcpyAsync(d[q], s[q], n * 8, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dst.tag, src.tag, n * 4, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dst.ev, src.ev, n, hipMemcpyDeviceToDevice, c->stream));
        return WC_OK;
    };
    unsigned long long steps0 = 0;
    HIPCHK(c, hipMemcpy(&steps0, c->d_steps, sizeof steps0, hipMemcpyDeviceToHost));
    // the events first: after the population is saved, every path restores it
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        return fail(c, WC_ERR_HIP, "hipEventCreate");
    }
    int rc = copy(c->s2, c->s);
    if (rc) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
    }
    const int tiles[] = {0, 1, 2, 4, 8};
    const int origTile = c->sortTile, origPeriod = c->sortPeriod;
    const bool origTiming = c->timing;
    c->timing = false;
    int keep = origTile;
    float best = 1e30f;
    for (int trial = -1; trial < 5 && !rc; ++trial) {       // trial -1 warms up
        const int tile = tiles[trial < 0 ? 1 : trial];
        if ((rc = copy(c->s, c->s2))) break;
        c->sortTile = tile;
        c->sortPeriod = 1;
        c->sortCount = 0;
        c->permValid = false;
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(c, WC_ERR_HIP, "hipEventRecord"); break; }
        if ((rc = wc_move_interval(c, cfg, start, end, slot0, slot1, path))) break;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(c, WC_ERR_HIP, "autotune trial with sort tile %d failed", tile);
            break;
        }
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(c, WC_ERR_HIP, "hipEventElapsedTime");
            break;
        }
        if (trial >= 0 && ms < best) {
            best = ms;
            keep = tile;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // the saved population goes back after a failed trial too
    const int rcRestore = copy(c->s, c->s2);
    if (!rc) rc = rcRestore;
    c->timing = origTiming;
    c->sortTile = rc ? origTile : keep;
    c->sortPeriod = rc ? origPeriod : 1;
    c->sortCount = 0;
    c->permValid = false;
    HIPCHK(c, hipMemcpyAsync(c->d_steps, &steps0, sizeof steps0, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (rc) return rc;
    if (best_tile) *best_tile = keep;
    if (best_ms) *best_ms = best;
    return WC_OK;
}

int wc_get_events(wc_ctx *c, uint8_t *events, size_t n) {
    if (!c || (n && !events)) return WC_ERR_INVALID;
    if (n > c->n) return fail(c, WC_ERR_INVALID, "asked for %zu particles, context holds %zu", n, c->n);
    if (!n) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(events, c->s.ev, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

int wc_move_interval(wc_ctx *c, const wc_config_data *cfg, double start, double end, int slot0, int slot1,
                     int path) {
    if (!c || !cfg) return WC_ERR_INVALID;
    int rc = check_supported(c, cfg);
    if (rc) return rc;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    const int ns = (int)c->slots.size();
    if (slot0 < 0 || slot0 >= ns || slot1 < 0 || slot1 >= ns || !c->slots[slot0].filled || !c->slots[slot1].filled)
        return fail(c, WC_ERR_INVALID, "forcing slots %d/%d not uploaded", slot0, slot1);
    if (path != WC_PATH_PARITY && path != WC_PATH_FAST) return fail(c, WC_ERR_INVALID, "unknown path %d", path);
    if (!std::isfinite(start) || !std::isfinite(end)) return fail(c, WC_ERR_INVALID, "non-finite interval");
    MoveParams P;
    if (!make_params(cfg, start, end, c->s_w, c->s_rho, c->eta, c->xi, P)) return WC_OK;
    if (c->n == 0) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const auto &a = c->slots[slot0], &b = c->slots[slot1];
    GridPtrs G{c->d_h, c->d_depth, c->d_dx, c->d_dy, nullptr, nullptr, c->d_hf, c->d_md, nullptr, c->d_depthf};
    RecordPair<float> F{a.f[0], b.f[0], a.f[1], b.f[1], a.f[2], b.f[2], a.f[3], b.f[3], a.f[4], b.f[4]};
    int rc2 = sort_population(c);
    if (rc2) return rc2;
    const uint32_t *perm = c->sortTile > 0 ? c->d_perm : nullptr;
    // the pack (timed as group 4) and the move kernel (group 0)
    for (int stage : {kStagePack, kStageMove}) {
        timing_begin(c);
        if (path == WC_PATH_PARITY)
            HIPCHK(c, launch_move_parity_soa(P, F, G, c->d_pack, c->d_metric, c->s, (int)c->n, c->d_steps,
                                             c->trackEvents, perm, stage, c->stream));
        else
            HIPCHK(c, launch_move_fast_soa(c->store, P, F, c->d_pack, G, c->s, (int)c->n, c->d_steps,
                                           c->trackEvents, perm, stage, c->lds, c->stream));
        if (stage == kStageMove) {
            // an upload into either slot now waits for this launch
            HIPCHK(c, hipEventRecord(c->slots[slot0].released, c->stream));
            HIPCHK(c, hipEventRecord(c->slots[slot1].released, c->stream));
        }
        const int trc = timing_end(c, stage == kStageMove ? 0 : 4);
        if (trc) return trc;
    }
    return WC_OK;
}

int wc_remove_dead(wc_ctx *c, size_t *remaining) {
    if (!c) return WC_ERR_INVALID;
    if (c->n == 0) {
        if (remaining) *remaining = 0;
        return WC_OK;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const bool keepPerm = c->permValid && c->sortTile > 0;
    HIPCHK(c, launch_compact(c->s, c->s2, (int)c->n, c->d_blockCounts, c->d_total, keepPerm ? c->d_remap : nullptr,
                             c->d_scanSums, c->stream));
    unsigned long long total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, c->d_total, sizeof total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (keepPerm) {             // the cell-tile order of the survivors, at their new indices
        HIPCHK(c, launch_perm_remap(c->d_perm, c->d_remap, (int)c->n, c->d_perm2, c->d_blockCounts, c->d_total,
                                    c->d_scanSums, c->stream));
        std::swap(c->d_perm, c->d_perm2);
    } else {
        c->permValid = false;
    }
    std::swap(c->s, c->s2);
    c->n = (size_t)total;
    if (remaining) *remaining = c->n;
    return WC_OK;
}

int wc_grid(wc_ctx *c, uint32_t *counts_out, float *conc_out, int mask_output) {
    if (!c) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t plane = (size_t)c->eta * c->xi, total = plane * c->s_rho;
    timing_begin(c);
    HIPCHK(c, launch_fill_u32(c->d_counts, total, 0u, c->stream));
    // population order: the LDS aggregation works as well, and the reads stay contiguous
    HIPCHK(c, launch_grid_count_soa(c->s, (int)c->n, c->s_rho, c->eta, c->xi, c->d_counts, nullptr, c->stream));
    // the reference's float frame (counts, land masked) stays on the device
    HIPCHK(c, launch_counts_to_conc(c->d_counts, c->d_mask, mask_output, c->s_rho, plane, c->d_conc, c->stream));
    int rc = timing_end(c, 1);
    if (rc) return rc;
    if (counts_out)
        HIPCHK(c, hipMemcpyAsync(counts_out, c->d_counts, total * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if (conc_out)
        HIPCHK(c, hipMemcpyAsync(conc_out, c->d_conc, total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (counts_out || conc_out) HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

int wc_emit(wc_ctx *c, const wc_config_data *cfg, const wc_source *sources, int n_sources, double start,
            double end, size_t *emitted) {
    if (emitted) *emitted = 0;
    if (!c || !cfg || n_sources < 0 || (n_sources && !sources)) return WC_ERR_INVALID;
    struct Emission {
        double time;
        int source;
    };
    std::vector<Emission> em;
    std::vector<double> t;
    for (int q = 0; q < n_sources; ++q) {
        if (!emission_times(sources[q].mode, sources[q].value, sources[q].start, sources[q].end, start, end, t))
            return fail(c, WC_ERR_INVALID, "invalid emission schedule of source %d", q);
        for (double x : t) em.push_back({x, q});
    }
    std::stable_sort(em.begin(), em.end(), [](const Emission &a, const Emission &b) { return a.time < b.time; });
    if (em.empty()) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    uint64_t firstId = 0;
    if (c->n) {                                               // Source.cpp:27: last id + 1
        HIPCHK(c, hipMemcpy(&firstId, c->s.id + (c->n - 1), 8, hipMemcpyDeviceToHost));
        firstId += 1;
    }
    const size_t m = em.size();
    int rc = reserve(c, c->n + m);
    if (rc) return rc;
    std::vector<double> times(m), pos(3 * (size_t)n_sources);
    std::vector<int> src(m);
    for (size_t q = 0; q < m; ++q) { times[q] = em[q].time; src[q] = em[q].source; }
    for (int q = 0; q < n_sources; ++q) {
        pos[3 * q] = sources[q].k; pos[3 * q + 1] = sources[q].j; pos[3 * q + 2] = sources[q].i;
    }
    DeviceTemp t_times, t_src, t_pos;
    HIPCHK(c, hipMalloc(&t_times.p, m * 8));
    HIPCHK(c, hipMalloc(&t_src.p, m * 4));
    HIPCHK(c, hipMalloc(&t_pos.p, pos.size() * 8));
    double *d_times = t_times.as<double>(), *d_pos = t_pos.as<double>();
    int *d_src = t_src.as<int>();
    HIPCHK(c, hipMemcpy(d_times, times.data(), m * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d_src, src.data(), m * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(d_pos, pos.data(), pos.size() * 8, hipMemcpyHostToDevice));
    const EmitDrift drift{cfg->driftObjectType, cfg->driftSide, cfg->leewayRandomSide ? 1 : 0,
                          cfg->leewayRightSideProbability};
    if (c->permValid) HIPCHK(c, launch_perm_fill(c->d_perm, c->n, c->n + m, c->stream));   // emitted last
    HIPCHK(c, launch_emit(c->s, c->n, (int)m, d_times, d_src, d_pos, firstId, cfg->randomSeed,
                          cfg->randomSources ? 1 : 0, drift, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n += m;
    if (emitted) *emitted = m;
    return WC_OK;
}

int wc_run_window(wc_ctx *c, const wc_config_data *cfg, const wc_window *win, const wc_source *sources,
                  int n_sources, int path, float *conc, int mask_output, wc_run_stats *stats) {
    if (!c || !cfg || !win || !win->oceanTime || win->T < 1) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    int rc = check_supported(c, cfg);
    if (rc) return rc;
    const int T = win->T, dir = cfg->trackingDirection;
    for (int t = 1; t < T; ++t)                               // ROMSAdapter.cpp:197-201
        if (!(win->oceanTime[t] > win->oceanTime[t - 1]))
            return fail(c, WC_ERR_INVALID, "ocean time must be strictly chronological (record %d)", t);
    const size_t plane = (size_t)c->eta * c->xi, frame = plane * c->s_rho;
    const size_t nf[5] = {plane, plane * c->s_rho, plane * c->s_rho, plane * c->s_w, plane * c->s_w};
    const float *fields[5] = {win->zeta, win->u, win->v, win->w, win->akt};
    if (conc) std::memset(conc, 0, frame * T * sizeof(float));
    wc_run_stats st{};
    const bool wasTiming = c->timing;
    c->timing = true;
    rc = timing_collect(c);
    if (rc) return rc;
    const double ms0 = c->ms[0];
    uint64_t steps0 = 0;
    rc = wc_step_count(c, &steps0);
    if (rc) return rc;
    // record rec -> slot rec % ring.  With three or more slots the record
    // after next is uploaded while the move runs (the copy waits only for the
    // move that last used its slot); with two, uploads and moves alternate.
    const int ring = c->slots.size() >= 3 ? 3 : 2;
    auto ensure = [&](int rec) -> int {
        auto &s = c->slots[rec % ring];
        if (s.filled && s.record == rec) return WC_OK;
        const float *src[5];
        for (int f = 0; f < 5; ++f) src[f] = fields[f] + nf[f] * rec;
        int r = wc_upload_record(c, rec % ring, src[0], src[1], src[2], src[3], src[4]);
        if (r) return r;
        s.record = rec;
        return WC_OK;
    };
    const auto t0 = std::chrono::steady_clock::now();
    const int first = dir < 0 ? T - 1 : 0;                   // Wacomm.cpp:200-206
    for (int idx = first; idx >= 0 && idx < T && idx + dir >= 0 && idx + dir < T; idx += dir) {
        const double start = win->oceanTime[idx], end = win->oceanTime[idx + dir];
        const double ck = cfg->restartCheckpoint;
        if (std::isfinite(ck) && std::isnan(restart_elapsed(start, end, ck))) continue;   // activeInterval
        if (dir > 0 && (!std::isfinite(ck) || start >= ck) && n_sources > 0) {            // :264-269
            rc = wc_emit(c, cfg, sources, n_sources, start, end, nullptr);
            if (rc) break;
        }
        if ((rc = ensure(idx)) || (rc = ensure(idx + dir))) break;
        rc = wc_move_interval(c, cfg, start, end, idx % ring, (idx + dir) % ring, path);
        if (rc) break;
        const int ahead = idx + 2 * dir;
        if (ring == 3 && ahead >= 0 && ahead < T && (rc = ensure(ahead))) break;
        size_t remaining = 0;
        if ((rc = wc_remove_dead(c, &remaining))) break;     // :710-713
        if ((rc = wc_grid(c, nullptr, conc ? conc + frame * (idx + dir) : nullptr, mask_output))) break;   // :722-739
        st.intervals += 1;
    }
    st.solver_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) {
        c->timing = wasTiming;
        return rc;
    }
    rc = timing_collect(c);
    c->timing = wasTiming;
    if (rc) return rc;
    uint64_t steps1 = 0;
    if ((rc = wc_step_count(c, &steps1))) return rc;
    st.move_ms = c->ms[0] - ms0;
    st.particle_steps = steps1 - steps0;
    st.particles = c->n;
    if (stats) *stats = st;
    return WC_OK;
}

// ---- batched ensemble ------------------------------------------------------

namespace {

constexpr int kEnsMaxSteps = 4096;

int alloc_ens_soa(wc_ctx *c, EnsSoA &s, size_t n) {
    HIPCHK(c, hipMalloc(&s.id, n * 8));
    HIPCHK(c, hipMalloc(&s.cell, n * 4));
    HIPCHK(c, hipMalloc(&s.fj, n * 4));
    HIPCHK(c, hipMalloc(&s.fi, n * 4));
    HIPCHK(c, hipMalloc(&s.k, n * 4));
    HIPCHK(c, hipMalloc(&s.health, n * 4));
    HIPCHK(c, hipMalloc(&s.age, n * 8));
    HIPCHK(c, hipMalloc(&s.time, n * 8));
    HIPCHK(c, hipMalloc(&s.scen, n));
    return WC_OK;
}

int copy_ens_soa(wc_ctx *c, const EnsSoA &dst, const EnsSoA &src, size_t n) {
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    HIPCHK(c, hipMemcpyAsync(dst.id, src.id, n * 8, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.cell, src.cell, n * 4, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.fj, src.fj, n * 4, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.fi, src.fi, n * 4, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.k, src.k, n * 4, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.health, src.health, n * 4, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.age, src.age, n * 8, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.time, src.time, n * 8, k, c->stream));
    HIPCHK(c, hipMemcpyAsync(dst.scen, src.scen, n, k, c->stream));
    return WC_OK;
}

// the cell-tile sort's counters for its smallest tile
constexpr int kMinSortTile = 2;
size_t ens_max_bins(const wc_ctx *c) {
    const size_t tx = (c->xi + kMinSortTile - 1) / kMinSortTile, ty = (c->eta + kMinSortTile - 1) / kMinSortTile;
    return tx * ty + 1;
}

EnsLaunch ens_launch(wc_ctx *c, bool surface) {
    EnsLaunch L{};
    L.steps = c->ens.d_steps;
    L.stepIdx = c->ens.d_stepIdx;
    L.packBuf = c->d_pack;
    L.G = GridPtrs{c->d_h, c->d_depth, c->d_dx, c->d_dy, nullptr, nullptr, c->d_hf, c->d_md, nullptr, c->d_depthf};
    L.s = c->ens.s;
    L.n = (int)c->ens.n;
    L.sigma = c->ens.d_sigma;
    L.nScen = c->ens.nScen;
    L.surface = surface ? c->ens.d_surface : nullptr;
    L.stepCounter = c->ens.d_count;
    L.s_w = c->s_w; L.s_rho = c->s_rho; L.eta = c->eta; L.xi = c->xi;
    L.block = c->ens.block; L.ppt = c->ens.ppt; L.occ = c->ens.occ;
    L.sortTile = c->ens.sortTile;
    L.sortPeriod = c->ens.sortPeriod;
    L.tmp = c->ens.tmp;
    L.bins = c->ens.d_bins;
    L.ldsChunk = c->ens.ldsChunk;
    L.ldsHalo = c->ens.ldsHalo;
    return L;
}

// Device step array for intervals [times[q], times[q+1]] with records in
// slots[q], slots[q+1]; inactive intervals (restart checkpoint) take no
// substeps.  Resets the device step counter.
int upload_ens_steps(wc_ctx *c, const wc_config_data *cfg, const double *times, const int *slots, int n) {
    std::vector<EnsStep> st(n);
    for (int q = 0; q < n; ++q) {
        if (!std::isfinite(times[q]) || !std::isfinite(times[q + 1]))
            return fail(c, WC_ERR_INVALID, "non-finite interval %d", q);
        const int a = slots[q], b = slots[q + 1], ns = (int)c->slots.size();
        if (a < 0 || a >= ns || b < 0 || b >= ns || !c->slots[a].filled || !c->slots[b].filled)
            return fail(c, WC_ERR_INVALID, "forcing slots %d/%d not uploaded", a, b);
        MoveParams &P = st[q].P;
        if (!make_params(cfg, times[q], times[q + 1], c->s_w, c->s_rho, c->eta, c->xi, P))
            P.elapsed0 = P.intervalLength;                    // inactive: no substeps
        const auto &A = c->slots[a], &B = c->slots[b];
        st[q].F = RecordPair<float>{A.f[0], B.f[0], A.f[1], B.f[1], A.f[2], B.f[2], A.f[3], B.f[3], A.f[4], B.f[4]};
    }
    HIPCHK(c, hipMemcpyAsync(c->ens.d_steps, st.data(), n * sizeof(EnsStep), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_fill_u32((uint32_t *)c->ens.d_stepIdx, 1, 0u, c->stream));
    // the host vector must outlive the copy
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

}  // namespace

int wc_ens_set(wc_ctx *c, const wc_particle_data *p, const uint8_t *scenario, size_t n, const double *sigma,
               int n_scenarios) {
    if (!c || (n && (!p || !scenario)) || !sigma) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    if (c->store > WC_STORE_BF16)
        return fail(c, WC_ERR_UNSUPPORTED, "the ensemble reads F32, F16 or BF16 storage (storage type %d)", c->store);
    if (n_scenarios < 1 || n_scenarios > 256) return fail(c, WC_ERR_INVALID, "1..256 scenarios");
    if (n > (size_t)kMaxParticles) return fail(c, WC_ERR_INVALID, "at most %d ensemble particles", kMaxParticles);
    const size_t plane = (size_t)c->eta * c->xi;
    if ((size_t)n_scenarios * plane > 0xFFFFFFFFull) return fail(c, WC_ERR_INVALID, "surface grids too large");
    for (size_t q = 0; q < n; ++q) {
        if (scenario[q] >= n_scenarios) return fail(c, WC_ERR_INVALID, "particle %zu: scenario %d", q, scenario[q]);
        if (p[q].driftObjectType != 0)
            return fail(c, WC_ERR_UNSUPPORTED, "ensemble particles are passive tracers (particle %zu has type %u)", q,
                        (unsigned)p[q].driftObjectType);
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_ens(c);
    auto &E = c->ens;
    const size_t cap = n ? n : 1;
    int rc = alloc_ens_soa(c, E.s, cap);
    if (!rc) rc = alloc_ens_soa(c, E.save, cap);
    if (!rc) rc = alloc_ens_soa(c, E.tmp, cap);
    if (rc) { free_ens(c); return rc; }
    HIPCHK(c, hipMalloc(&E.d_bins, ens_max_bins(c) * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&E.d_sigma, n_scenarios * sizeof(float)));
    HIPCHK(c, hipMalloc(&E.d_surface, (size_t)n_scenarios * plane * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&E.d_steps, kEnsMaxSteps * sizeof(EnsStep)));
    HIPCHK(c, hipMalloc(&E.d_stepIdx, sizeof(int)));
    HIPCHK(c, hipMalloc(&E.d_count, sizeof(unsigned long long)));
    // host split into cell + FP32 offsets; a position outside the
    // interpolation cells gets an invalid cell (it ends at its first substep,
    // as validInterpolationCell fails, NumericalHelpers.hpp:12-14)
    std::vector<uint64_t> id(n);
    std::vector<uint32_t> cell(n);
    std::vector<float> fj(n), fi(n), k(n), health(n), sg(n_scenarios);
    std::vector<double> age(n), time(n);
    for (size_t q = 0; q < n; ++q) {
        const wc_particle_data &d = p[q];
        const double jfl = std::floor(d.j), ifl = std::floor(d.i);
        const bool inside = jfl >= 0 && ifl >= 0 && jfl + 1 < c->eta && ifl + 1 < c->xi;
        float oj = inside ? (float)(d.j - jfl) : 0.f, oi = inside ? (float)(d.i - ifl) : 0.f;
        long jc = (long)jfl, ic = (long)ifl;
        if (oj >= 1.0f) { oj -= 1.0f; ++jc; }
        if (oi >= 1.0f) { oi -= 1.0f; ++ic; }
        id[q] = d.id;
        cell[q] = inside ? (uint32_t)(jc * c->xi + ic) : 0xFFFFFFFFu;
        fj[q] = oj; fi[q] = oi;
        k[q] = (float)d.k;
        health[q] = (float)d.health;
        age[q] = d.age;
        time[q] = d.time;
    }
    for (int q = 0; q < n_scenarios; ++q) sg[q] = (float)sigma[q];
    if (n) {
        HIPCHK(c, hipMemcpy(E.s.id, id.data(), n * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.cell, cell.data(), n * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.fj, fj.data(), n * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.fi, fi.data(), n * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.k, k.data(), n * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.health, health.data(), n * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.age, age.data(), n * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.time, time.data(), n * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(E.s.scen, scenario, n, hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemcpy(E.d_sigma, sg.data(), n_scenarios * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, launch_fill_u32(E.d_surface, (size_t)n_scenarios * plane, 0u, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    E.n = n;
    E.nScen = n_scenarios;
    E.have = true;
    return WC_OK;
}

int wc_ens_release(wc_ctx *c, const wc_ens_source *sources, int n_sources, const double *sigma, int n_scenarios,
                   uint64_t seed) {
    if (!c || (n_sources && !sources) || !sigma || n_sources < 0) return WC_ERR_INVALID;
    if (!c->haveGrid) return fail(c, WC_ERR_INVALID, "wc_set_grid first");
    if (c->store > WC_STORE_BF16)
        return fail(c, WC_ERR_UNSUPPORTED, "the ensemble reads F32, F16 or BF16 storage (storage type %d)", c->store);
    if (n_scenarios < 1 || n_scenarios > 256) return fail(c, WC_ERR_INVALID, "1..256 scenarios");
    const size_t plane = (size_t)c->eta * c->xi;
    if ((size_t)n_scenarios * plane > 0xFFFFFFFFull) return fail(c, WC_ERR_INVALID, "surface grids too large");
    std::vector<EnsSource> src(n_sources);
    std::vector<unsigned long long> first(n_sources + 1, 0);
    for (int q = 0; q < n_sources; ++q) {
        const wc_ens_source &S = sources[q];
        if (S.scenario < 0 || S.scenario >= n_scenarios || S.sub_scenario >= n_scenarios ||
            !std::isfinite(S.j) || !std::isfinite(S.i) || !std::isfinite(S.k) || !(S.spread >= 0) ||
            !std::isfinite(S.time))
            return fail(c, WC_ERR_INVALID, "invalid ensemble source %d", q);
        src[q] = EnsSource{S.k, S.j, S.i, S.spread, S.time, S.scenario, S.sub_scenario, S.sub_every};
        // checked before the sum, so a huge count cannot wrap it
        if (S.count > (unsigned long long)kMaxParticles - first[q])
            return fail(c, WC_ERR_INVALID, "at most %d ensemble particles (source %d)", kMaxParticles, q);
        first[q + 1] = first[q] + S.count;
    }
    const unsigned long long n = first[n_sources];
    if (n > (unsigned long long)kMaxParticles)
        return fail(c, WC_ERR_INVALID, "at most %d ensemble particles", kMaxParticles);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_ens(c);
    auto &E = c->ens;
    const size_t cap = n ? n : 1;
    int rc = alloc_ens_soa(c, E.s, cap);
    if (!rc) rc = alloc_ens_soa(c, E.save, cap);
    if (!rc) rc = alloc_ens_soa(c, E.tmp, cap);
    if (rc) { free_ens(c); return rc; }
    HIPCHK(c, hipMalloc(&E.d_bins, ens_max_bins(c) * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&E.d_sigma, n_scenarios * sizeof(float)));
    HIPCHK(c, hipMalloc(&E.d_surface, (size_t)n_scenarios * plane * sizeof(uint32_t)));
    HIPCHK(c, hipMalloc(&E.d_steps, kEnsMaxSteps * sizeof(EnsStep)));
    HIPCHK(c, hipMalloc(&E.d_stepIdx, sizeof(int)));
    HIPCHK(c, hipMalloc(&E.d_count, sizeof(unsigned long long)));
    std::vector<float> sg(n_scenarios);
    for (int q = 0; q < n_scenarios; ++q) sg[q] = (float)sigma[q];
    HIPCHK(c, hipMemcpy(E.d_sigma, sg.data(), n_scenarios * sizeof(float), hipMemcpyHostToDevice));
    DeviceTemp t_src, t_first;
    HIPCHK(c, hipMalloc(&t_src.p, (n_sources ? n_sources : 1) * sizeof(EnsSource)));
    HIPCHK(c, hipMalloc(&t_first.p, (n_sources + 1) * sizeof(unsigned long long)));
    EnsSource *d_src = t_src.as<EnsSource>();
    unsigned long long *d_first = t_first.as<unsigned long long>();
    hipError_t e = hipSuccess;
    if (n_sources) e = hipMemcpy(d_src, src.data(), n_sources * sizeof(EnsSource), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d_first, first.data(), (n_sources + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_ens_release(d_src, d_first, n_sources, (long long)n, c->eta, c->xi, seed, E.s, c->stream);
    if (e == hipSuccess) e = launch_fill_u32(E.d_surface, (size_t)n_scenarios * plane, 0u, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, WC_ERR_HIP, "ensemble release: %s", hipGetErrorString(e));
    E.n = n;
    E.nScen = n_scenarios;
    E.have = true;
    return WC_OK;
}

int wc_ens_set_launch(wc_ctx *c, int block, int ppt, int occ) {
    if (!c) return WC_ERR_INVALID;
    if ((block != 64 && block != 128 && block != 256) || ppt < 1 || ppt > 64 || (occ != 0 && occ != 3 && occ != 4))
        return fail(c, WC_ERR_INVALID, "launch block=%d ppt=%d occ=%d not offered", block, ppt, occ);
    if (occ == 4 && c->ens.ldsChunk > 0)
        return fail(c, WC_ERR_INVALID, "the LDS-tiled move runs at occupancy hint 0 or 3");
    free_ens_graph(c);
    c->ens.block = block; c->ens.ppt = ppt; c->ens.occ = occ;
    return WC_OK;
}

int wc_ens_set_sort(wc_ctx *c, int tile, int period) {
    if (!c) return WC_ERR_INVALID;
    if ((tile != 0 && (tile < kMinSortTile || tile > 64)) || period < 1)
        return fail(c, WC_ERR_INVALID, "sort tile %d (0 or %d..64), period %d (>= 1)", tile, kMinSortTile, period);
    free_ens_graph(c);
    c->ens.sortTile = tile;
    c->ens.sortPeriod = period;
    return WC_OK;
}

int wc_ens_set_lds(wc_ctx *c, int chunk, int halo) {
    if (!c) return WC_ERR_INVALID;
    if (chunk < 0 || chunk > 1024 || halo < 0 || halo > 16)
        return fail(c, WC_ERR_INVALID, "LDS chunk %d (0..1024), halo %d (0..16)", chunk, halo);
    if (chunk > 0 && c->ens.occ == 4)
        return fail(c, WC_ERR_INVALID, "the LDS-tiled move runs at occupancy hint 0 or 3");
    free_ens_graph(c);
    c->ens.ldsChunk = chunk;
    c->ens.ldsHalo = halo;
    return WC_OK;
}

int wc_ens_get_lds(wc_ctx *c, int *chunk, int *halo) {
    if (!c) return WC_ERR_INVALID;
    if (chunk) *chunk = c->ens.ldsChunk;
    if (halo) *halo = c->ens.ldsHalo;
    return WC_OK;
}

int wc_ens_get_sort(wc_ctx *c, int *tile, int *period) {
    if (!c) return WC_ERR_INVALID;
    if (tile) *tile = c->ens.sortTile;
    if (period) *period = c->ens.sortPeriod;
    return WC_OK;
}

int wc_ens_get_launch(wc_ctx *c, int *block, int *ppt, int *occ) {
    if (!c) return WC_ERR_INVALID;
    if (block) *block = c->ens.block;
    if (ppt) *ppt = c->ens.ppt;
    if (occ) *occ = c->ens.occ;
    return WC_OK;
}

int wc_ens_autotune(wc_ctx *c, const wc_config_data *cfg, double start, double end, int slot0, int slot1,
                    double *best_ms) {
    if (!c || !cfg) return WC_ERR_INVALID;
    if (!c->ens.have) return fail(c, WC_ERR_INVALID, "wc_ens_set first");
    int rc = check_supported(c, cfg);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    auto &E = c->ens;
    const double times[2] = {start, end};
    const int slots[2] = {slot0, slot1};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        return fail(c, WC_ERR_HIP, "hipEventCreate");
    }
    rc = copy_ens_soa(c, E.save, E.s, E.n);                    // the population to restore
    if (rc) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
    }
    // candidates: the plain move over block x particles per work-item x
    // occupancy hint, and the LDS-tiled move (256 x 1) over occupancy hint x
    // chunk; each with every sort tile
    struct Cand {
        int block, ppt, occ, tile, chunk;
    };
    std::vector<Cand> cands;
    for (int tile : {0, 2, 4, 8}) {
        for (int occ : {0, 3, 4})
            for (int ppt : {1, 2, 4})
                for (int block : {64, 128, 256}) cands.push_back(Cand{block, ppt, occ, tile, 0});
        for (int occ : {0, 3})
            for (int chunk : {10, 30}) cands.push_back(Cand{256, 1, occ, tile, chunk});
    }
    double best = INFINITY;
    const Cand orig{E.block, E.ppt, E.occ, E.sortTile, E.ldsChunk};
    const int origPeriod = E.sortPeriod;
    Cand keep = orig;
    E.sortPeriod = 1;
    const int nc = (int)cands.size();
    for (int trial = -1; trial < nc && !rc; ++trial) {
        const Cand &t = cands[trial < 0 ? nc - 1 : trial];       // trial -1 warms up the last candidate
        E.block = t.block; E.ppt = t.ppt; E.occ = t.occ; E.sortTile = t.tile; E.ldsChunk = t.chunk;
        if ((rc = copy_ens_soa(c, E.s, E.save, E.n))) break;
        if ((rc = upload_ens_steps(c, cfg, times, slots, 1))) break;
        const EnsLaunch L = ens_launch(c, false);
        if (hipEventRecord(e0, c->stream) != hipSuccess || launch_ens_step(c->store, L, c->stream) != hipSuccess ||
            hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(c, WC_ERR_HIP, "autotune launch block=%d ppt=%d occ=%d tile=%d chunk=%d failed", E.block,
                      E.ppt, E.occ, E.sortTile, E.ldsChunk);
            break;
        }
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(c, WC_ERR_HIP, "hipEventElapsedTime");
            break;
        }
        if (trial >= 0 && ms < best) {
            best = ms;
            keep = t;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    for (int q = 0; q < 2 && !rc; ++q)
        if (hipEventRecord(c->slots[slots[q]].released, c->stream) != hipSuccess)
            rc = fail(c, WC_ERR_HIP, "hipEventRecord failed");
    // the saved population goes back after a failed trial too
    const int rcRestore = copy_ens_soa(c, E.s, E.save, E.n);
    if (!rc) rc = rcRestore;
    if (rc) {                                   // leave the caller's launch as it was
        E.block = orig.block; E.ppt = orig.ppt; E.occ = orig.occ; E.sortTile = orig.tile;
        E.ldsChunk = orig.chunk; E.sortPeriod = origPeriod;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (rc) return rc;
    free_ens_graph(c);
    E.block = keep.block; E.ppt = keep.ppt; E.occ = keep.occ; E.sortTile = keep.tile; E.ldsChunk = keep.chunk;
    E.tuned = true;
    if (best_ms) *best_ms = best;
    return WC_OK;
}

int wc_ens_run(wc_ctx *c, const wc_config_data *cfg, const double *times, const int *slots, int n_intervals,
               int use_graph, int surface, wc_ens_stats *stats) {
    if (!c || !cfg || !times || !slots || n_intervals < 0) return WC_ERR_INVALID;
    if (!c->ens.have) return fail(c, WC_ERR_INVALID, "wc_ens_set first");
    if (n_intervals > kEnsMaxSteps) return fail(c, WC_ERR_INVALID, "at most %d intervals per call", kEnsMaxSteps);
    int rc = check_supported(c, cfg);
    if (rc) return rc;
    if (stats) *stats = wc_ens_stats{};
    if (n_intervals == 0) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    auto &E = c->ens;
    rc = upload_ens_steps(c, cfg, times, slots, n_intervals);
    if (rc) return rc;
    HIPCHK(c, launch_fill_u32(reinterpret_cast<uint32_t *>(E.d_count), 2, 0u, c->stream));
    const EnsLaunch L = ens_launch(c, surface != 0);
    const bool wantSurface = surface != 0;
    if (use_graph) {
        // the captured step depends on the launch shape and on whether it grids
        if (E.exec && (E.graphSurface != wantSurface)) free_ens_graph(c);
        if (!E.exec) {
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t le = launch_ens_step(c->store, L, c->stream);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(c->stream, &g);
            if (le != hipSuccess || ce != hipSuccess) {
                if (g) hipGraphDestroy(g);
                return fail(c, WC_ERR_HIP, "capturing the ensemble step failed: %s / %s", hipGetErrorString(le),
                            hipGetErrorString(ce));
            }
            const hipError_t ie = hipGraphInstantiate(&E.exec, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                E.exec = nullptr;
                hipGraphDestroy(g);
                return fail(c, WC_ERR_HIP, "instantiating the ensemble step failed: %s", hipGetErrorString(ie));
            }
            E.graph = g;
            E.graphSurface = wantSurface;
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        return fail(c, WC_ERR_HIP, "hipEventCreate");
    }
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t he = hipEventRecord(e0, c->stream);
    for (int q = 0; q < n_intervals && he == hipSuccess; ++q)
        he = use_graph ? hipGraphLaunch(E.exec, c->stream) : launch_ens_step(c->store, L, c->stream);
    // the slots' last reader is now enqueued (an upload into them waits for it)
    for (int q = 0; q <= n_intervals && he == hipSuccess; ++q) he = hipEventRecord(c->slots[slots[q]].released, c->stream);
    if (he == hipSuccess) he = hipEventRecord(e1, c->stream);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    const auto t1 = std::chrono::steady_clock::now();
    float ms = 0;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (he != hipSuccess) return fail(c, WC_ERR_HIP, "ensemble run: %s", hipGetErrorString(he));
    unsigned long long count = 0;
    HIPCHK(c, hipMemcpy(&count, E.d_count, sizeof count, hipMemcpyDeviceToHost));
    if (stats) {
        stats->device_ms = ms;
        stats->wall_s = std::chrono::duration<double>(t1 - t0).count();
        stats->particle_steps = count;
        stats->intervals = n_intervals;
    }
    return WC_OK;
}

int wc_ens_surface(wc_ctx *c, uint32_t *counts) {
    if (!c || !counts) return WC_ERR_INVALID;
    if (!c->ens.have) return fail(c, WC_ERR_INVALID, "wc_ens_set first");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t cells = (size_t)c->ens.nScen * c->eta * c->xi;
    HIPCHK(c, hipMemcpyAsync(counts, c->ens.d_surface, cells * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

int wc_ens_get(wc_ctx *c, wc_particle_data *out, uint8_t *scenario, size_t n) {
    if (!c || (n && !out)) return WC_ERR_INVALID;
    if (!c->ens.have) return fail(c, WC_ERR_INVALID, "wc_ens_set first");
    if (n > c->ens.n) return fail(c, WC_ERR_INVALID, "asked for %zu particles, the ensemble holds %zu", n, c->ens.n);
    if (!n) return WC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const EnsSoA &S = c->ens.s;
    std::vector<uint64_t> id(n);
    std::vector<uint32_t> cell(n);
    std::vector<float> fj(n), fi(n), k(n), health(n);
    std::vector<double> age(n), time(n);
    std::vector<uint8_t> sc(n);
    const hipMemcpyKind D2H = hipMemcpyDeviceToHost;
    HIPCHK(c, hipMemcpyAsync(id.data(), S.id, n * 8, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(cell.data(), S.cell, n * 4, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(fj.data(), S.fj, n * 4, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(fi.data(), S.fi, n * 4, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(k.data(), S.k, n * 4, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(health.data(), S.health, n * 4, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(age.data(), S.age, n * 8, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(time.data(), S.time, n * 8, D2H, c->stream));
    HIPCHK(c, hipMemcpyAsync(sc.data(), S.scen, n, D2H, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // id order: the cell-tile sort leaves the population in an order of its own
    std::vector<size_t> ord(n);
    bool dense = c->ens.n == n;
    std::vector<char> seen(dense ? n : 0, 0);
    for (size_t q = 0; q < n && dense; ++q) {
        if (id[q] >= n || seen[id[q]]) dense = false;
        else { seen[id[q]] = 1; ord[id[q]] = q; }
    }
    if (!dense) {
        for (size_t q = 0; q < n; ++q) ord[q] = q;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return id[a] < id[b]; });
    }
    for (size_t r = 0; r < n; ++r) {
        const size_t q = ord[r];
        wc_particle_data d{};
        d.id = id[q];
        const bool valid = cell[q] != 0xFFFFFFFFu;
        const uint32_t jc = valid ? cell[q] / (uint32_t)c->xi : 0, ic = valid ? cell[q] - jc * (uint32_t)c->xi : 0;
        d.j = valid ? (double)jc + (double)fj[q] : NAN;
        d.i = valid ? (double)ic + (double)fi[q] : NAN;
        d.k = k[q];
        d.health = health[q];
        d.age = age[q];
        d.time = time[q];
        out[r] = d;
        if (scenario) scenario[r] = sc[q];
    }
    return WC_OK;
}

int wc_set_timing(wc_ctx *c, int enabled) {
    if (!c) return WC_ERR_INVALID;
    c->timing = enabled != 0;
    return WC_OK;
}

int wc_kernel_time(wc_ctx *c, int kernel, double *total_ms, long *launches) {
    if (!c || kernel < 0 || kernel >= kTimedKernels) return WC_ERR_INVALID;
    int rc = timing_collect(c);
    if (rc) return rc;
    if (total_ms) *total_ms = c->ms[kernel];
    if (launches) *launches = c->launches[kernel];
    return WC_OK;
}

int wc_reset_kernel_time(wc_ctx *c) {
    if (!c) return WC_ERR_INVALID;
    int rc = timing_collect(c);
    if (rc) return rc;
    for (int k = 0; k < kTimedKernels; ++k) {
        c->ms[k] = 0;
        c->launches[k] = 0;
    }
    return WC_OK;
}

int wc_step_count(wc_ctx *c, uint64_t *steps) {
    if (!c || !steps) return WC_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    unsigned long long v = 0;
    HIPCHK(c, hipMemcpyAsync(&v, c->d_steps, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *steps = v;
    return WC_OK;
}

int wc_reset_step_count(wc_ctx *c) {
    if (!c) return WC_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemsetAsync(c->d_steps, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WC_OK;
}

}  // extern "C"
