/*
 * wacomm_hip.h -- C ABI of libwacomm_hip.so, the MI355X (gfx950) particle-move
 * and concentration-gridding backend for WaComM++.
 *
 * Two layers:
 *   1. Drop-in entry points with the reference's own argument lists
 *      (wc_hip_move_particles replaces cudaMoveParticle; wc_hip_concentration
 *      replaces Concentration::evaluate).  All pointers are device pointers, as
 *      in the reference's CUDA call.
 *   2. The native context API: the library owns the particle state (SoA) and a
 *      ring of forcing records on the device (two slots = the reference's
 *      bounded current/adjacent pair), and runs the interval loop of
 *      Wacomm::run (move -> remove dead -> grid) without host round trips.
 *
 * Conventions: every function returns a wc_status (never exits the process);
 * wc_last_error() returns the message of the last failure on a context.  All
 * work is asynchronous on the context's stream unless stated; wc_sync() waits.
 * Calls on one context must be serialised by the caller; one context per
 * device.  Plain C types only (no HIP or torch types in the signatures).
 */
#ifndef WACOMM_HIP_H
#define WACOMM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define WC_BOOL bool
extern "C" {
#else
#define WC_BOOL _Bool
#endif

/* ---------------------------------------------------------------------------
 * Shared data structures (byte-compatible with the reference)
 * ------------------------------------------------------------------------- */

/* observational_calibration_region, Config.hpp:25-34 */
typedef struct wc_calibration_region {
    double west, south, east, north;
    double windErrorStdDev, windErrorComponentCorrelation;
    double windErrorSpatialScale, windErrorTemporalScale;
} wc_calibration_region;

/* config_data, Config.hpp:36-72 (same field order, types and padding) */
typedef struct wc_config_data {
    WC_BOOL random;
    WC_BOOL randomSources;
    double deltat;
    double dti;
    double survprob;
    double tau0;
    double sv;
    double crid;
    double sigma;
    double shoreLimit;
    int upperClosure;
    int lowerClosure;
    int horizontalClosure;
    uint64_t randomSeed;
    int trackingDirection;
    int backwardDiffusion;
    int integrator;
    double restartCheckpoint;
    int driftModel;
    WC_BOOL leewayCoefficientEnsemble;
    double leewayResidualCorrelation;
    WC_BOOL leewayRandomSide;
    double leewayRightSideProbability;
    double leewayJibeProbabilityHourly;
    double windErrorStdDev;
    double windErrorComponentCorrelation;
    double windErrorSpatialScale;
    double windErrorTemporalScale;
    uint32_t observationalCalibrationRegionCount;
    wc_calibration_region observationalCalibrationRegions[16];
    uint16_t driftObjectType;
    int8_t driftSide;
    WC_BOOL hasWind;
    double windU10;
    double windV10;
} wc_config_data;

/* particle_data, Particle.hpp:31-41 (64 bytes) */
typedef struct wc_particle_data {
    uint64_t id;
    double k;
    double j;
    double i;
    double health;
    double age;
    double time;
    uint16_t driftObjectType;
    int8_t driftSide;
} wc_particle_data;

typedef enum wc_status {
    WC_OK = 0,
    WC_ERR_HIP = 1,          /* a HIP runtime call or kernel launch failed */
    WC_ERR_INVALID = 2,      /* bad argument (sizes, slots, null pointers) */
    WC_ERR_UNSUPPORTED = 3,  /* configuration the device path does not implement
                                (driftModel==1, jibing): use the CPU backend */
    WC_ERR_NOMEM = 4,        /* device allocation failed */
    WC_ERR_NO_DEVICE = 5,    /* no HIP device, or an invalid device index */
    WC_ERR_IO = 6            /* a file could not be written */
} wc_status;

/* Closure and integrator codes, Config.hpp:201-207 */
enum { WC_CLOSURE_CONSTRAINT = 1, WC_CLOSURE_KILL = 2, WC_CLOSURE_REFLECTION = 3 };
enum { WC_INTEGRATOR_EULER = 1, WC_INTEGRATOR_MIDPOINT = 2 };

/* Forcing storage on the device (the move always reads FP32 values from the
 * host; the context converts once per record upload).
 *   WC_STORE_F32, WC_STORE_F16, WC_STORE_BF16: every value rounded to the type.
 *   WC_STORE_BF16_TILE, WC_STORE_F16_TILE: u, v and w stored as 16-bit
 *     residuals from a mean per tile of 4 x 4 cells (per level, field and
 *     record: the mean of the tile's non-zero values, itself rounded to the
 *     16-bit type, so land zeros decode to exactly 0), so the 8 (BF16) or 11
 *     (FP16) significant bits resolve the variation inside a tile instead of
 *     the whole value; AKt and zeta are rounded to the 16-bit type.  Each
 *     tile keeps its own copy of the 5 x 5 nodes its cells' stencils read, so
 *     a stencil needs one mean.  Mixed-precision path only.
 *   WC_STORE_I16_TILE: the same tiles, each residual an int16 multiple of a
 *     power-of-two scale per tile (one per field and W level, shared by the
 *     two records; the mean an integer multiple of it), so every node of a
 *     tile is resolved to 2^-16 of the tile's largest deviation from its mean
 *     (block floating point); AKt as int16 multiples of its own tile scale,
 *     zeta kept in FP32.  Land zeros decode to exactly 0.  Mixed-precision
 *     path only. */
enum {
    WC_STORE_F32 = 0,
    WC_STORE_F16 = 1,
    WC_STORE_BF16 = 2,
    WC_STORE_BF16_TILE = 3,
    WC_STORE_F16_TILE = 4,
    WC_STORE_I16_TILE = 5
};

/* Closure events of a particle's trajectory, OR-ed per particle by the move
 * kernels while event tracking is on (wc_set_event_tracking):
 *   WC_EVENT_SHORE     the candidate position was on land (h + zeta <=
 *                      shoreLimit) and the horizontal closure acted
 *                      (Particle.cpp:753-771);
 *   WC_EVENT_DOMAIN    the candidate left the interpolation domain
 *                      (Particle.cpp:777-791);
 *   WC_EVENT_FALLBACK  the midpoint X* was invalid or on land and the step
 *                      fell back to Euler (Particle.cpp:603-641);
 *   WC_EVENT_VERTICAL  the surface or bottom closure acted (Particle.cpp:
 *                      134-176, 665-704);
 *   WC_EVENT_SHALLOW   the water column at the start position was at or
 *                      below shoreLimit (Particle.cpp:311-338);
 *   WC_EVENT_COAST     a substep started in a cell with a land corner (h = 0,
 *                      as ROMSAdapter.cpp:244-252 sets it on land), where the
 *                      interpolated velocity falls to the land's zero. */
enum {
    WC_EVENT_SHORE = 1,
    WC_EVENT_DOMAIN = 2,
    WC_EVENT_FALLBACK = 4,
    WC_EVENT_VERTICAL = 8,
    WC_EVENT_SHALLOW = 16,
    WC_EVENT_COAST = 32
};

/* Move paths of wc_move_interval:
 *   WC_PATH_PARITY: FP64 state and the reference's rounding points
 *                   (positions match Particle::move to ~1e-12 cells); reads
 *                   FP32 forcing.
 *   WC_PATH_FAST:   integer cell + FP32 offset state, FP32 interpolation of
 *                   F32/F16/BF16 forcing, Philox4x32-10 diffusion draws keyed
 *                   on (seed, id, interval, substep). */
enum { WC_PATH_PARITY = 0, WC_PATH_FAST = 1 };

/* ---------------------------------------------------------------------------
 * 1. Drop-in entry points
 * ------------------------------------------------------------------------- */

/* Replaces cudaMoveParticle (cuda/kernel.h:7, cuda/kernel.cu:370-382): moves
 * particleCount particles in place over the forcing interval that starts at
 * times[timeIndex], exactly as Particle::move (Particle.cpp:71-824) does.
 * The haversine cell sizes (Particle.cpp:544-574) come from a table the
 * library computes on the host once per lon/lat grid and keeps per device
 * (a content hash of lon/lat, taken on the device at every call, detects a
 * new grid), so results equal the CPU path's bit for bit without diffusion;
 * calls on one device serialise on the host while they enqueue.
 * config, particles, times and all field arrays are DEVICE pointers laid out
 * as the reference's (index4 = ((t*levels + k+levels-1)*eta + j)*xi + i,
 * kernel.cu:63-69).  mask, numThread, numGPU are unused (as in the reference);
 * windU10..stokesV must be NULL or are ignored for passive particles.
 * stream is a hipStream_t (NULL = default stream).  Asynchronous: results are
 * valid after the stream is synchronised.  Non-passive particles move
 * passively (driftModel 0) and, with leewayJibeProbabilityHourly > 0, flip
 * their side as Particle.cpp:805-813 does.  Returns WC_ERR_UNSUPPORTED when
 * config->driftModel==1. */
int wc_hip_move_particles(wc_config_data *config, wc_particle_data *particles, int timeIndex, int oceanTime,
                          int sW, int sRho, int eta, int xi, double *times, double *mask, double *lonRad,
                          double *latRad, double *depthIntervals, double *h, float *zeta, float *u, float *v,
                          float *w, float *akt, float *windU10, float *windV10, float *stokesU, float *stokesV,
                          int particleCount, int numThread, int numGPU, void *stream);

/* Replaces Concentration::evaluate (Concentration.hpp:16-44) for one frame:
 * conc[s_rho][eta][xi] (device, float, storage level k+s_rho-1) += 1 for every
 * alive particle whose (int)round(k), (int)round(j), (int)round(i) lie in the
 * grid.  Counts are accumulated exactly in uint32 and added once per cell,
 * which equals the reference's repeated +1.0f below 2^24 per cell.
 * particles is a device pointer. */
int wc_hip_concentration(const wc_particle_data *particles, int particleCount, int s_rho, int eta, int xi,
                         float *conc, void *stream);

/* Allocation-free gridding.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  The result of wc_grid's float frame for a device
 * particle_data array: conc[s_rho][eta][xi] (device) is overwritten with the
 * count of alive particles per cell (Concentration.hpp:16-44), as float
 * (saturating at 2^24), and 1e37 on land cells (mask != 1) when mask_output is
 * nonzero (Wacomm.cpp:724-739; mask: device [eta][xi] doubles, required
 * then).  The exact counts accumulate in a static device buffer of
 * WC_GRID_NOALLOC_MAX_CELLS uint32 (64 MiB) that each device the library
 * runs on holds until the process ends; s_rho * eta * xi above it returns
 * WC_ERR_INVALID.  Calls no hipMalloc, hipFree or allocating routine; the
 * kernels run asynchronously on stream (a hipStream_t, NULL = default).
 * Every call on a device shares its buffer (each call clears the cells it
 * uses first), so calls that overlap on one device race. */
#define WC_GRID_NOALLOC_MAX_CELLS 16777216
int wc_grid_noalloc(const wc_particle_data *particles, int particleCount, int s_rho, int eta, int xi,
                    const double *mask, int mask_output, float *conc, void *stream);

/* ---------------------------------------------------------------------------
 * 2. Native context API
 * ------------------------------------------------------------------------- */

typedef struct wc_ctx wc_ctx;

/* Creates a context on HIP device `device` with its own stream (default
 * priority).  Replaces the per-interval setup of Wacomm.cpp:386-482. */
int wc_ctx_create(int device, wc_ctx **out);
int wc_ctx_destroy(wc_ctx *ctx);
/* Message of the last failure on ctx (empty when none); valid until the next call. */
const char *wc_last_error(const wc_ctx *ctx);
int wc_sync(wc_ctx *ctx);

/* Static grid of the forcing window (OceanModelAdapter.hpp:20-34), host
 * pointers copied to the device.  lonRad, latRad, h, mask: [eta][xi];
 * depthIntervals: s_w-1 values (entry kI+s_w-2 = depthIntervals(kI) of the
 * reference, ROMSAdapter.cpp:233-235).  Also precomputes each cell's
 * haversine sizes (Particle.cpp:544-574) in FP64 on the host.
 * n_slots = resident forcing records (>= 2; 2 = the current/adjacent pair of
 * WacommPlusPlus.cpp:98-161).  store = WC_STORE_*. */
int wc_set_grid(wc_ctx *ctx, int s_w, int s_rho, int eta, int xi, const double *lonRad, const double *latRad,
                const double *h, const double *mask, const double *depthIntervals, int n_slots, int store);

/* Uploads one forcing record (host FP32 arrays: zeta [eta][xi], u, v
 * [s_rho][eta][xi], w, akt [s_w][eta][xi]) into ring slot `slot`; slots hold
 * FP32 and each interval's pack converts the record pair to the context's
 * storage type.  Replaces the whole-window re-upload of Wacomm.cpp:417-477
 * by one record per interval.  The copy runs on the context's copy stream and
 * waits only for enqueued moves that read this slot, so with three or more
 * slots the record after next uploads while the current move runs.  Returns
 * when the copy is complete: the host arrays may be reused at once. */
int wc_upload_record(wc_ctx *ctx, int slot, const float *zeta, const float *u, const float *v, const float *w,
                     const float *akt);

/* Copies ring slot `slot` back to the host as the move reads it (FP32, the
 * layout of wc_upload_record); a NULL field is skipped.  Replaces reads of
 * OceanModelAdapter::Zeta/U/V/W/AKT (OceanModelAdapter.hpp:20-34), e.g. the
 * assertions of tests/ROMSAdapterTest.cpp:67-74. */
int wc_get_record(wc_ctx *ctx, int slot, float *zeta, float *u, float *v, float *w, float *akt);

/* The static grid of a ROMS file, as ROMSAdapter::process reads it
 * (ROMSAdapter.cpp:56-166).  All arrays are host pointers, row-major:
 * lon_rho, lat_rho (degrees), h, mask_rho, angle (radians): [eta_rho][xi_rho];
 * mask_u: [eta_u][xi_u]; mask_v: [eta_v][xi_v]; s_w: the s_rho+1 sigma values
 * of the W levels in file order (bottom, -1, first).  Masks hold 0 or 1.
 * earth_relative: u and v carry standard_name eastward/northward_sea_water_
 * velocity, so no rotation (angle may then be NULL).  missing_*: the
 * variable's _FillValue / missing_value attributes (at most
 * WC_ROMS_MAX_MISSING each); the NetCDF float and double fill values are
 * always treated as missing too (ROMSAdapter.cpp:29-43).  The attribute and
 * units checks on the file's metadata belong to the file reader. */
#define WC_ROMS_MAX_MISSING 14
typedef struct wc_roms_grid {
    int s_rho, eta_rho, xi_rho, eta_u, xi_u, eta_v, xi_v;
    const double *s_w;
    const double *lon_rho, *lat_rho, *h;
    const double *mask_rho, *mask_u, *mask_v;
    const double *angle;
    int earth_relative;
    const double *missing_u, *missing_v, *missing_angle;
    int n_missing_u, n_missing_v, n_missing_angle;
    const double *s_rho_values;   /* optional (NULL): s_rho sigma values, bottom first, for
                                     wc_check_roms_grid */
} wc_roms_grid;

/* Replaces the static part of ROMSAdapter::process (ROMSAdapter.cpp:
 * 146-151, 228-275): checks the masks (0/1 only) and the angle at wet rho
 * points (finite, not missing), then calls wc_set_grid with lonRad = latRad
 * = 0.0174533 * degrees, h zeroed on land, depthIntervals(k) = s_w(k) -
 * s_w(k-1), mask = mask_rho.  The u and v grids must not exceed the rho grid
 * (ROMS staggering: eta_u = eta_rho, xi_u = xi_rho-1, eta_v = eta_rho-1,
 * xi_v = xi_rho).  Returns WC_ERR_INVALID where the reference throws. */
int wc_set_roms_grid(wc_ctx *ctx, const wc_roms_grid *grid, int n_slots, int store);

/* Replaces the per-record part of ROMSAdapter::process (ROMSAdapter.cpp:
 * 152-166, 254-300, 309-368, 448-466) on the device, writing ring slot
 * `slot`: host FP32 zeta [eta_rho][xi_rho], u [s_rho][eta_u][xi_u],
 * v [s_rho][eta_v][xi_v], w, akt [s_rho+1][eta_rho][xi_rho].  u and v are
 * averaged over wet faces onto wet rho points (float), rotated by the grid
 * angle into east/north (double, rounded to float) unless earth_relative;
 * w, akt are zeroed on land; zeta is zeroed on land and where >= 1e36.
 * Bit-identical to the reference's arithmetic.  Returns WC_ERR_INVALID (slot
 * left unfilled) for a non-finite or missing u/v at a wet face or a rotated
 * velocity outside float range, where the reference throws. */
int wc_upload_roms_record(wc_ctx *ctx, int slot, const float *zeta, const float *u, const float *v,
                          const float *w, const float *akt);

/* The geographic grid the particle records need: lat, lon (degrees)
 * [eta][xi] and the s_w W-level sigma values bottom first (the
 * OceanModelAdapter's Lat, Lon and SW), host pointers; the bathymetry is the
 * grid's h.  wc_set_roms_grid sets it from lon_rho, lat_rho and s_w. */
int wc_set_geo(wc_ctx *ctx, const double *lat, const double *lon, const double *s_w);

/* Particles::netcdfRecords (Particles.cpp:115-147) on the device: the
 * context's particles with j >= 0, i >= 0, k <= 0 that
 * OceanModelAdapter::kji2deplatlon (OceanModelAdapter.cpp:271-321) places
 * inside the interpolation domain, in population order, with their depth
 * (negative down), lat and lon by bilinear interpolation; bit-identical to
 * the reference's FP64 arithmetic.  *count receives the number of records;
 * with out non-NULL the non-NULL fields of out (host arrays of capacity
 * entries) receive them.  WC_ERR_INVALID when capacity is too small. */
typedef struct wc_records {
    uint64_t *id;
    double *lat, *lon, *depth, *k, *j, *i, *health, *age, *time;
    uint32_t *object_type;   /* driftObjectType (Particles.cpp:144) */
    int32_t *drift_side;     /* driftSide (Particles.cpp:145) */
} wc_records;
int wc_particle_records(wc_ctx *ctx, const wc_records *out, size_t capacity, size_t *count);

/* The grid checks of OceanModelAdapter::appendBoundaryRecord
 * (OceanModelAdapter.cpp:199-214) for the next file of a run, against the
 * grid of the last wc_set_roms_grid: same dimensions; s_rho values (when both
 * were given) and s_w values within 1e-12; lon_rho, lat_rho within 1e-10;
 * h (zeroed on land) within 1e-8; identical mask_rho.  Returns WC_OK, or
 * WC_ERR_INVALID with the reference's message. */
int wc_check_roms_grid(wc_ctx *ctx, const wc_roms_grid *next);

/* The duplicate-boundary check of appendBoundaryRecord
 * (OceanModelAdapter.cpp:216-233) on the device: compares slots a and b
 * under float != (a NaN always conflicts).  *conflict = 0 when they agree,
 * else the class the reference throws for, at the first conflicting point in
 * its (j, i) loop order: 1 surface elevation, 2 horizontal velocity,
 * 3 vertical fields (w, AKt).  *point (optional) = that point j*xi+i, or -1. */
int wc_compare_records(wc_ctx *ctx, int slot_a, int slot_b, int *conflict, long long *point);

/* Particle state: host AoS in and out (Particles.hpp:17 holds particle_data).
 * wc_set_particles replaces the device population; wc_append_particles adds
 * to it (emission).  A context holds at most 2,147,483,392 particles (the
 * kernels index with 32-bit integers; the drop-in entry points take the same
 * limit for particleCount).  Every field travels, the drift object type and side
 * included: with driftModel 0 a non-passive particle moves as a passive one
 * (Particle.cpp:380-454 act only under driftModel 1), and the hourly side
 * flip of leewayJibeProbabilityHourly > 0 (Particle.cpp:805-813) runs on
 * both move paths with the reference's keyed draw. */
int wc_set_particles(wc_ctx *ctx, const wc_particle_data *particles, size_t n);
int wc_append_particles(wc_ctx *ctx, const wc_particle_data *particles, size_t n);
int wc_get_particles(wc_ctx *ctx, wc_particle_data *particles, size_t n);
int wc_particle_count(const wc_ctx *ctx, size_t *n);

/* Closure-event tracking (no counterpart in the reference; a diagnostic of
 * which trajectories touched a closure, e.g. to separate particles held on
 * coast node lines from the rest when two move paths are compared).  While
 * enabled, both move paths OR WC_EVENT_* bits into a per-particle byte that
 * travels with the particle through compaction; particles set, appended or
 * emitted start at 0.  wc_clear_events zeroes them; wc_get_events copies the
 * first n (population order) to the host. */
int wc_set_event_tracking(wc_ctx *ctx, int enabled);

/* Cell-tile order of the moves (no counterpart in the reference): with
 * tile > 0, wc_move_interval visits the particles tile by tile of tile x
 * tile cells, so the lanes of a wavefront gather forcing from neighbouring
 * cells (wc_grid keeps the population order: its per-block LDS hash already
 * aggregates equal cells, and the permuted reads measured slower).  The order is a permutation kept beside the population,
 * which itself stays in the reference's order (every result, and
 * wc_get_particles, is unchanged); it is recomputed every period-th move and
 * follows compaction, release and appends.  tile 0 (the default) turns it
 * off.  wc_autotune_sort times one move over [start, end] with tile 0, 1, 2,
 * 4 and 8 on the current population (restored afterwards, as is the step
 * count), keeps the fastest with period 1, and returns it in *tile (and its
 * time in *best_ms); the sort's own time is timed as kernel group 3. */
int wc_set_sort(wc_ctx *ctx, int tile, int period);
int wc_get_sort(wc_ctx *ctx, int *tile, int *period);
int wc_autotune_sort(wc_ctx *ctx, const wc_config_data *config, double start, double end, int slot0, int slot1,
                     int path, int *tile, double *best_ms);
/* LDS-staged forcing for the mixed-precision move (no counterpart in the
 * reference): with chunk > 0, WC_PATH_FAST moves of the plain packed storage
 * types (WC_STORE_BF16, WC_STORE_F16, WC_STORE_F32) run the interval in
 * chunks of `chunk` substeps, and before each chunk every workgroup copies
 * the interval pack of the cells its particles occupy, widened by `halo`
 * cells (shrunk until the box fits in 47 KiB), into LDS; stencils inside the
 * box read LDS, any other reads global memory.  The values read, and so
 * every result, are those of the plain move.  Other storage types, the
 * drift-side jibe and event tracking use the plain move.  chunk 0 (the
 * default) turns it off; chunk <= 1024, halo <= 16. */
int wc_set_lds(wc_ctx *ctx, int chunk, int halo);
int wc_get_lds(wc_ctx *ctx, int *chunk, int *halo);
int wc_clear_events(wc_ctx *ctx);
int wc_get_events(wc_ctx *ctx, uint8_t *events, size_t n);

/* One forcing interval of Particle::move for every particle (the loop of
 * Wacomm.cpp:501-525).  config is a HOST pointer.  slot0 holds the record at
 * the interval start, slot1 the record at its end (the same slot when the
 * window has no next record, Particle.cpp:183-184).  path = WC_PATH_*.
 * Asynchronous on the context stream. */
int wc_move_interval(wc_ctx *ctx, const wc_config_data *config, double intervalStart, double intervalEnd,
                     int slot0, int slot1, int path);

/* Order-preserving removal of particles with health <= 0 (Wacomm.cpp:710-713).
 * Synchronous: *remaining receives the new count. */
int wc_remove_dead(wc_ctx *ctx, size_t *remaining);

/* Concentration::evaluate (Concentration.hpp:16-44) into the context's device
 * count frame [s_rho][eta][xi] (uint32, zeroed first), then the reference's
 * float frame (Wacomm.cpp:722-739) on the device: counts as float, and 1e37
 * on land cells (mask != 1) when mask_output is nonzero.  counts_out and
 * conc_out, when not NULL, receive host copies of the two (synchronous). */
int wc_grid(wc_ctx *ctx, uint32_t *counts_out, float *conc_out, int mask_output);

/* A release point (Source.hpp) and its schedule (SourceEmission.hpp:12-16).
 * start/end are seconds on the forcing time axis (end < 0: open-ended). */
typedef struct wc_source {
    double k, j, i;
    int mode;          /* 0 uniform_rate, 1 forcing_interval_batch, 2 single_pulse */
    double value;      /* particles per hour / per forcing record / in the pulse */
    double start, end;
} wc_source;

/* Replaces Sources::emit (Sources.cpp:233-248), Source::emit (Source.cpp:25-44)
 * and sourceEmissionTimes (SourceEmission.hpp:38-79): appends the particles
 * released in [intervalStart, intervalEnd) in time order, ids continuing from
 * the last particle's id (0 when empty); with config->randomSources each
 * coordinate gets 0.25*normal(seed, id, SOURCE_POSITION_RANDOM_INTERVAL, 0, c)
 * (SourceEmission.hpp:18-23).  Each particle takes config->driftObjectType
 * and config->driftSide, or with config->leewayRandomSide the side
 * sampleDriftSide draws (RIGHT when uniform(seed, id, LEEWAY_SIDE_RANDOM_INTERVAL,
 * 0, 0) < leewayRightSideProbability, else LEFT; Source.cpp:38-42,
 * DriftModel.hpp:146-154).  *emitted receives the count.  Returns
 * WC_ERR_INVALID for an invalid schedule (where the reference throws). */
int wc_emit(wc_ctx *ctx, const wc_config_data *config, const wc_source *sources, int n_sources,
            double intervalStart, double intervalEnd, size_t *emitted);

/* A forcing window on the host (oceanmodel_data, OceanModelAdapter.hpp:20-34):
 * T records, zeta [T][eta][xi], u/v [T][s_rho][eta][xi], w/akt [T][s_w][eta][xi],
 * oceanTime [T] strictly increasing (else WC_ERR_INVALID, as
 * ROMSAdapter.cpp:197-201 throws).  The grid must match the context's
 * wc_set_grid. */
typedef struct wc_window {
    int T;
    const double *oceanTime;
    const float *zeta, *u, *v, *w, *akt;
} wc_window;

typedef struct wc_run_stats {
    double solver_s;          /* wall time of the interval loop */
    double move_ms;           /* device time of the move kernels */
    uint64_t particle_steps;  /* live substeps taken */
    size_t particles;         /* particles alive at the end */
    int intervals;            /* intervals processed */
} wc_run_stats;

/* The interval loop of Wacomm::run (Wacomm.cpp:200-770) for one window:
 * for each interval in the tracking direction that is active under the
 * restart checkpoint (NumericalHelpers.hpp:54-60): emit (forward tracking,
 * emitAtIntervalStart), move, remove dead, grid into frame idx+direction.
 * Records stream through a two-slot ring (one upload per interval).  conc,
 * when not NULL, is the host [T][s_rho][eta][xi] float array of
 * Wacomm.cpp:172-186 (zeroed here), with land cells = 1e37 when mask_output. */
int wc_run_window(wc_ctx *ctx, const wc_config_data *config, const wc_window *window, const wc_source *sources,
                  int n_sources, int path, float *conc, int mask_output, wc_run_stats *stats);

/* Accumulated device time of the move kernels launched since the last reset,
 * measured with HIP events on the context stream around each launch (enable
 * with wc_set_timing).  kernel: 0 = move, 1 = grid, 2 = ROMS normalisation,
 * 3 = the cell-tile sort of wc_set_sort, 4 = the interval pack that precedes
 * each move (the record pair rearranged for the move's gathers). */
int wc_set_timing(wc_ctx *ctx, int enabled);
int wc_kernel_time(wc_ctx *ctx, int kernel, double *total_ms, long *launches);
int wc_reset_kernel_time(wc_ctx *ctx);

/* Live particle-substeps taken by the moves since the last reset (counted on
 * the device; the numerator of particle-steps/s).  Synchronous. */
int wc_step_count(wc_ctx *ctx, uint64_t *steps);
int wc_reset_step_count(wc_ctx *ctx);

/* ---------------------------------------------------------------------------
 * 3. Host utilities (no device needed)
 * ------------------------------------------------------------------------- */

/* Library version string. */
const char *wc_version(void);

/* Haversine cell sizes of Particle.cpp:544-574 for every cell of an
 * [eta][xi] lon/lat grid (radians): dx toward (j,i+1), dy toward (j+1,i);
 * the last row and column hold 0.  The context computes its table with this. */
void wc_cell_metric_table(int eta, int xi, const double *lonRad, const double *latRad, double *dx, double *dy);

/* Philox4x32-10 block (the generator of WC_PATH_FAST's diffusion draws):
 * out = philox(ctr[4], key[2]).  No reference counterpart; exposed so the
 * generator can be checked against the published known-answer vectors. */
void wc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---------------------------------------------------------------------------
 * 3. Batched ensemble (no counterpart in the reference: many release
 *    scenarios, each with its own source, horizontal diffusion and release
 *    offset, moved as one population on the mixed-precision path)
 * ------------------------------------------------------------------------- */

/* Loads the ensemble population on the context's grid (wc_set_grid first;
 * the storage type of wc_set_grid is the forcing precision).  particles[q]
 * belongs to scenario scenario[q] < n_scenarios <= 256; sigma[s] is scenario
 * s's horizontal diffusion scale (sqrt(2 Kh dti)); a particle's emission
 * time is its release (late activation, Particle.cpp:211-215).  State is
 * compact: integer cell + FP32 offsets and level, FP32 health, FP64 age and
 * emission time.  Per-scenario surface-layer count grids
 * [n_scenarios][eta][xi] (uint32) are allocated here.  Ensemble particles
 * are passive tracers: a non-passive one is WC_ERR_UNSUPPORTED. */
int wc_ens_set(wc_ctx *ctx, const wc_particle_data *particles, const uint8_t *scenario, size_t n,
               const double *sigma, int n_scenarios);

/* A release point of a device-generated ensemble population. */
typedef struct wc_ens_source {
    double k, j, i;          /* release point (fractional indices) */
    double spread;           /* std dev, in cells, of each particle's offset in j and i */
    double time;             /* emission time (s) */
    uint64_t count;          /* particles */
    int scenario;            /* scenario of the particles */
    int sub_scenario;        /* scenario of every sub_every-th particle of this source, or -1 */
    uint32_t sub_every;      /* 0: no subset */
} wc_ens_source;

/* As wc_ens_set, but the population is generated on the device: particle q
 * (id q, sources in order) starts at its source's point offset by spread *
 * N(0, 1) in j and i (Philox4x32-10 keyed on seed and q); every sub_every-th
 * particle of a source belongs to sub_scenario (an id-subsample that moves
 * exactly as in the full population, since every draw is keyed by id). */
int wc_ens_release(wc_ctx *ctx, const wc_ens_source *sources, int n_sources, const double *sigma, int n_scenarios,
                   uint64_t seed);

/* Launch shape of the ensemble move: block in {64, 128, 256} work-items,
 * ppt particles per work-item (1..64), occupancy hint occ in {0 (none),
 * 3, 4} waves per SIMD.  wc_ens_autotune picks them by timing. */
int wc_ens_set_launch(wc_ctx *ctx, int block, int ppt, int occ);
int wc_ens_get_launch(wc_ctx *ctx, int *block, int *ppt, int *occ);
int wc_ens_get_sort(wc_ctx *ctx, int *tile, int *period);
/* Stages the forcing around each workgroup's particles in LDS: the interval
 * runs in chunks of `chunk` substeps, and before each chunk every workgroup
 * of 256 particles copies the packed forcing of its particles' bounding box
 * (widened by `halo` cells and one level each way) into LDS; stencils inside
 * the box read LDS, the others global memory.  The results are those of the
 * plain move, bit for bit.  chunk 0 (the default) turns it off; halo 0..16,
 * default 2.  Needs occupancy hint 0 or 3. */
int wc_ens_set_lds(wc_ctx *ctx, int chunk, int halo);
int wc_ens_get_lds(wc_ctx *ctx, int *chunk, int *halo);

/* Reorders the ensemble population by the tile (tile x tile cells) of its cell
 * before the move of every period-th step, so a wave's lanes gather from
 * neighbouring cells; tile 0 never sorts (the default).  Results are per
 * particle and do not depend on the order. */
int wc_ens_set_sort(wc_ctx *ctx, int tile, int period);

/* Times every offered launch on one interval [start, end] (records in slot0,
 * slot1) over the current population, which is restored afterwards, and
 * keeps the fastest (124 trials): the plain move over block x particles per
 * work-item x occupancy hint, and the LDS-tiled move (occupancy hint 0 or 3,
 * chunk 10 or 30 substeps, halo kept), each with sort tile none, 2, 4 or 8
 * (sorted every step).  *best_ms (optional) receives its time. */
int wc_ens_autotune(wc_ctx *ctx, const wc_config_data *config, double start, double end, int slot0, int slot1,
                    double *best_ms);

typedef struct wc_ens_stats {
    double device_ms;         /* device time of the whole run (events) */
    double wall_s;            /* host wall time of the launches and the final wait */
    uint64_t particle_steps;  /* live substeps taken */
    int intervals;
} wc_ens_stats;

/* Runs n_intervals forcing intervals: interval q spans times[q] ->
 * times[q+1] with records in slots[q], slots[q+1] (n_intervals+1 entries
 * each).  One step = pack the record pair, move the population, and when
 * `surface` is nonzero clear and fill the per-scenario surface counts, all
 * on the device reading the interval from a device array.  With use_graph
 * the step is captured once into a HIP graph (kept until the launch shape,
 * the grid or the population changes) and replayed per interval; otherwise
 * the same kernels are launched directly.  The two give identical results. */
int wc_ens_run(wc_ctx *ctx, const wc_config_data *config, const double *times, const int *slots, int n_intervals,
               int use_graph, int surface, wc_ens_stats *stats);

/* Per-scenario surface counts of the last step run with surface != 0:
 * counts[n_scenarios][eta][xi]. */
int wc_ens_surface(wc_ctx *ctx, uint32_t *counts);

/* The ensemble population as particle_data in id order (j = cell row +
 * offset, i = cell column + offset; NaN for a particle that started outside
 * the grid) and, when scenario is not NULL, each particle's scenario. */
int wc_ens_get(wc_ctx *ctx, wc_particle_data *particles, uint8_t *scenario, size_t n);

/* ---------------------------------------------------------------------------
 * 4. Learned concentration reconstruction (no counterpart in the reference:
 *    a convolutional network that maps a low-particle-count surface
 *    histogram to the high-count one, trained on pairs the method generates)
 * ------------------------------------------------------------------------- */

/* A 4-level U-Net with 32 base channels (32-64-128-256, skips by channel
 * concatenation, 2x2 average pooling, nearest upsampling, ReLU) and a
 * residual 1x1 head: out = x + head(dec1).  Forward and backward run on BF16
 * matrix cores with FP32 accumulation; Adam updates FP32 master parameters.
 * Parameters are FP32 in this order: per conv layer (enc1a, enc1b, enc2a,
 * enc2b, enc3a, enc3b, bott_a, bott_b, dec3a, dec3b, dec2a, dec2b, dec1a,
 * dec1b) weights [cout][3][3][cin] then biases [cout]; then the head's 32
 * weights and its bias.  cin of enc1a is 32 (the input channels first, then
 * zeros); the decoder's first convs take concat(skip, upsampled).  He-normal
 * initialisation from `seed`.  Images are [n][h][w] FP32 on the host (one
 * input channel), h and w multiples of 8. */
typedef struct wc_unet wc_unet;
int wc_unet_create(int device, uint64_t seed, wc_unet **out);
/* The same network with `channels` (1..32) input channels: images are then
 * [n][channels][h][w] FP32 (the first layer's input channels 0..channels-1;
 * the residual adds channel 0), e.g. a low-count histogram at several output
 * times.  wc_unet_create is channels = 1. */
int wc_unet_create_ex(int device, uint64_t seed, int channels, wc_unet **out);
int wc_unet_destroy(wc_unet *net);
const char *wc_unet_last_error(const wc_unet *net);
size_t wc_unet_param_count(const wc_unet *net);
int wc_unet_get_params(wc_unet *net, float *params);
/* Replaces the parameters and resets the optimiser state. */
int wc_unet_set_params(wc_unet *net, const float *params);
/* y = network(x); *device_ms (optional) = device time of the forward pass. */
int wc_unet_forward(wc_unet *net, const float *x, int n, int h, int w, float *y, double *device_ms);
/* Gradient of loss = mean (network(x) - t)^2 with respect to the parameters
 * (grad may be NULL); no update. */
int wc_unet_gradients(wc_unet *net, const float *x, const float *t, int n, int h, int w, float *grad, float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at learning rate lr on
 * that loss; *loss receives the loss before the step. */
int wc_unet_train_step(wc_unet *net, const float *x, const float *t, int n, int h, int w, float lr, float *loss);
/* The same two calls for the per-pixel weighted loss mean weight (network(x)
 * - t)^2, weight [n][h][w] (NULL: all ones, the calls above). */
int wc_unet_gradients_weighted(wc_unet *net, const float *x, const float *t, const float *weight, int n, int h,
                               int w, float *grad, float *loss);
int wc_unet_train_step_weighted(wc_unet *net, const float *x, const float *t, const float *weight, int n, int h,
                                int w, float lr, float *loss);
/* The gradient and training calls launch the forward and the backward pass
 * (about 55 kernels) as two HIP graphs, captured the first time a batch
 * shape (and loss kind) is seen and replayed afterwards (on = 1, the
 * default), or launch them one by one (on = 0).  Both run the same kernels
 * with the same arguments and give identical results.  The batch is staged
 * through pinned host memory the network owns; the calls return after the
 * device has finished. */
int wc_unet_set_graph(wc_unet *net, int on);

/* ---------------------------------------------------------------------------
 * 5. Output formats around the path
 * ------------------------------------------------------------------------- */

/* The concentration file of Wacomm::save (Wacomm.cpp:913-1020): dimensions
 * ocean_time, s_rho, eta_rho, eta_xi; variables ocean_time, s_rho, lat_rho,
 * lon_rho, mask_rho, h (double), zeta [T][eta][xi] and conc
 * [T][s_rho][eta][xi] (float, _FillValue 9.99999993e36) with the reference's
 * attributes, in the NetCDF classic format with 64-bit offsets (CDF-2).
 * zeta / conc may be NULL (written as _FillValue, as the reference leaves
 * conc unwritten when another rank writes it).  global_attributes: NULL or a
 * NULL-terminated list of name, value strings (the provenance attributes of
 * Provenance.hpp:19-29).  Written to <path>.partial, then renamed.
 * Host-only: needs no device.  Returns WC_ERR_IO when the file cannot be
 * written. */
typedef struct wc_conc_file {
    int T, s_rho, eta, xi;
    const double *ocean_time;     /* [T] */
    const double *s_rho_values;   /* [s_rho] */
    const double *lat_rho, *lon_rho, *mask_rho, *h;   /* [eta][xi] */
    const float *zeta;            /* [T][eta][xi] or NULL */
    const float *conc;            /* [T][s_rho][eta][xi] or NULL */
    const char *const *global_attributes;
} wc_conc_file;
int wc_write_concentration_nc(const char *path, const wc_conc_file *file);

/* A particle restart/history file, restart version 3, as
 * Particles::saveAsNetCDF writes it (Particles.cpp:150-264): the global
 * attributes global_attributes (provenance, Provenance.hpp:19-29; a
 * wacomm_random_key_scheme of ordered-splitmix64-v2 is added unless given),
 * wacomm_restart_version "3", tracking_direction, ocean_model,
 * checkpoint_time (double) and random_seed (uint64); dimensions particles (n)
 * and particle_time (1); variables particle_time, id (uint64), object_type
 * (uint), drift_side (int), lat, lon, depth, k, j, i, health, age, time
 * [particles][particle_time] (double) with the reference's attributes.  The
 * reference writes NetCDF-4 (HDF5); this file is the classic format with
 * 64-bit data (CDF-5), which holds uint64 and which the NetCDF library reads
 * through the same calls.  object_type / drift_side NULL are written as 0 (the
 * plain method's drift state).  An empty history (n = 0) keeps zero-length
 * variables.  Written to <path>.partial, then renamed.  Host-only.  Returns
 * WC_ERR_INVALID for a missing field or a global attribute that repeats
 * another, WC_ERR_IO when the file cannot be written. */
typedef struct wc_restart_file {
    size_t n;
    double particle_time;         /* the particle_time variable and checkpoint_time */
    uint64_t random_seed;
    int backward;                 /* tracking_direction: 1 "backward", 0 "forward" */
    const char *ocean_model;
    const char *const *global_attributes;
    const uint64_t *id;
    const uint32_t *object_type;
    const int32_t *drift_side;
    const double *lat, *lon, *depth, *k, *j, *i, *health, *age, *time;
} wc_restart_file;
int wc_write_restart_nc(const char *path, const wc_restart_file *file);

/* Particles::loadFromNetCDF (Particles.cpp:373-480) for NetCDF classic files
 * (CDF-1, CDF-2, CDF-5).  When the file has wacomm_restart_version, checks it
 * as the reference does against the configuration (cfg->trackingDirection,
 * randomSeed, driftModel == 1 for leeway) and ocean_model: version 2 or 3,
 * version 2 not with leeway, the tracking direction, random_seed,
 * wacomm_random_key_scheme and ocean_model when present; checkpoint_time, when
 * present, is stored in *checkpoint_time (the reference's
 * Config::RestartCheckpoint), which is otherwise left unchanged.  Reads id, k,
 * j, i, health, age, time (the first particle_time column) and, when present,
 * object_type and drift_side into out[0..n) in file order.  *count receives n;
 * out NULL only counts.  err (err_len bytes, may be NULL) receives the message.
 * Returns WC_ERR_INVALID with the reference's message where it throws,
 * WC_ERR_UNSUPPORTED for a NetCDF-4 (HDF5) file, WC_ERR_IO for a file that
 * cannot be read or is not NetCDF. */
int wc_read_restart_nc(const char *path, const wc_config_data *cfg, const char *ocean_model,
                       wc_particle_data *out, size_t capacity, size_t *count, double *checkpoint_time, char *err,
                       size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* WACOMM_HIP_H */
