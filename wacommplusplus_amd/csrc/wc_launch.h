// Host-side launchers of the device kernels (internal to libwacomm_hip.so).
#pragma once

#include <vector>

#include "wc_device.h"

namespace wc {

// Stages of a context-path move launch: the interval pack, the move kernel.
constexpr int kStagePack = 1, kStageMove = 2, kStagesAll = 3;

// Context path: packs the interval's record pair into packBuf
// (parity_pack_bytes) and moves the SoA population; metric = {dx, dy} table.
size_t parity_pack_bytes(int s_w, size_t plane);
// events: OR the closure events of each particle into s.ev.
// perm (optional): the order in which work-items take particles.
// stages: kStage* bits; the move stage reads the pack the pack stage wrote.
hipError_t launch_move_parity_soa(const MoveParams &P, const RecordPair<float> &F, const GridPtrs &G, void *packBuf,
                                  const double2 *metric, const SoA &s, int n, unsigned long long *stepCounter,
                                  bool events, const uint32_t *perm, int stages, hipStream_t stream);
// Drop-in path: the reference's array layout; metric = {dx, dy} table.
// With packBuf (parity_pack_bytes) the record pair is packed first and the
// move gathers from the pack; without, it gathers from the raw records.
hipError_t launch_move_parity_aos(const MoveParams &P, const RecordPair<float> &F, const GridPtrs &G,
                                  const double2 *metric, void *packBuf, wc_particle_data *p, int n,
                                  unsigned long long *stepCounter, hipStream_t stream);
// *hash += a content hash of lon[0..n) and lat[0..n) (zero it first).
hipError_t launch_grid_hash(const double *lon, const double *lat, size_t n, unsigned long long *hash,
                            hipStream_t stream);

// Mixed-precision path (store = WC_STORE_*): packs the interval's record pair
// into packBuf (fast_pack_bytes) and moves the particles.  G.idxf must point to
// the float2 {1/dx, 1/dy} table.
size_t fast_pack_bytes(int store, int s_w, int eta, int xi);
// LDS staging of the forcing for the move (plain packed storage, no jibe, no
// event tracking): chunk > 0 substeps per staged box, halo cells around it
struct LdsTiling {
    int chunk, halo;
};
constexpr int kMoveLdsBytes = 47 * 1024;     // per workgroup: three workgroups per CU
hipError_t launch_move_fast_soa(int store, const MoveParams &P, const RecordPair<float> &F, void *packBuf,
                                const GridPtrs &G, const SoA &s, int n, unsigned long long *stepCounter,
                                bool events, const uint32_t *perm, int stages, const LdsTiling &lt,
                                hipStream_t stream);

hipError_t launch_fill_u32(uint32_t *a, size_t n, uint32_t v, hipStream_t stream);
// perm (optional): visit the particles in that order
hipError_t launch_grid_count_soa(const SoA &s, int n, int s_rho, int eta, int xi, uint32_t *counts,
                                 const uint32_t *perm, hipStream_t stream);
hipError_t launch_grid_count_aos(const wc_particle_data *p, int n, int s_rho, int eta, int xi, uint32_t *counts,
                                 hipStream_t stream);
hipError_t launch_grid_count_aos_float(const wc_particle_data *p, int n, int s_rho, int eta, int xi, float *conc,
                                       hipStream_t stream);
hipError_t launch_counts_to_conc(const uint32_t *counts, const double *mask, int mask_output, int s_rho,
                                 size_t plane, float *conc, hipStream_t stream);
// remap (optional, n entries): each particle's new index, or 0xFFFFFFFF if it
// died.  scanSums (optional): kScanSumsCap entries of scratch for the
// multi-block scans (scans of up to 64M values; longer ones use one block).
constexpr size_t kScanSumsCap = 65536;
hipError_t launch_compact(const SoA &src, const SoA &dst, int n, uint32_t *blockCounts, unsigned long long *total,
                          uint32_t *remap, uint32_t *scanSums, hipStream_t stream);
// Cell-tile order: permOut = the population's indices grouped by tile of
// tile x tile cells (particles outside the grid last), read in the order
// permIn (nullptr: population order); bins: tiles + 1 counters.
hipError_t launch_sort_perm(const SoA &s, int n, int eta, int xi, int tile, const uint32_t *permIn,
                            uint32_t *permOut, uint32_t *bins, unsigned long long *total, uint32_t *scanSums,
                            hipStream_t stream);
// out = the entries of perm whose particle survived a compaction, mapped
// through remap, in perm order; total[0] = their number
hipError_t launch_perm_remap(const uint32_t *perm, const uint32_t *remap, int n, uint32_t *out,
                             uint32_t *blockCounts, unsigned long long *total, uint32_t *scanSums,
                             hipStream_t stream);
// perm[q] = q for q in [from, to)
hipError_t launch_perm_fill(uint32_t *perm, size_t from, size_t to, hipStream_t stream);

// Particle history records (Particles::netcdfRecords): the grid's lat, lon
// (degrees) and h [eta][xi], the s_w W-level sigma values bottom first.
struct GeoPtrs {
    const double *lat, *lon, *h, *sW;
    int eta, xi, s_w;
};
// Order-preserving: the particles of src inside the interpolation domain go
// to dst (state) and dep, lat, lon; total[0] receives their number.
hipError_t launch_particle_records(const SoA &src, int n, const GeoPtrs &g, const SoA &dst, double *dep,
                                   double *lat, double *lon, uint32_t *blockCounts, unsigned long long *total,
                                   hipStream_t stream);

hipError_t launch_emit(const SoA &s, size_t at, int n, const double *times, const int *src, const double *pos,
                       uint64_t firstId, uint64_t seed, int randomSources, EmitDrift drift, hipStream_t stream);
bool emission_times(int mode, double value, double sourceStart, double sourceEnd, double intervalStart,
                    double intervalEnd, std::vector<double> &out);

// ROMS record normalisation (wc_roms.hip).  missing: the variable's
// _FillValue / missing_value attributes plus the NetCDF float and double fill
// values.  flag bits: 1 bad u at a wet face, 2 bad v at a wet face, 4 a
// rotated velocity outside float range.
constexpr int kRomsMissing = 16;
constexpr float kNcFillFloat = 9.9692099683868690e+36f;     // netcdf.h NC_FILL_FLOAT
constexpr double kNcFillDouble = 9.9692099683868690e+36;    // netcdf.h NC_FILL_DOUBLE
struct RomsDims {
    int s_rho, eta, xi, eta_u, xi_u, eta_v, xi_v;
    int earth;
};
struct RomsMissing {
    double u[kRomsMissing], v[kRomsMissing];
    int nu, nv;
};
// face: per rho point, kFaceWet | kFaceU (u face (j,i) wet) | kFaceUWest (u
// face (j,i-1) wet) | kFaceV (v face (j,i) wet) | kFaceVSouth (v face (j-1,i)
// wet), faces outside their grid counted dry.
constexpr uint32_t kFaceWet = 1, kFaceU = 2, kFaceUWest = 4, kFaceV = 8, kFaceVSouth = 16;
hipError_t launch_roms_normalize(const RomsDims &d, const RomsMissing &miss, const uint8_t *face,
                                 const double2 *rot, const float *u, const float *v, float *U, float *V, float *W,
                                 float *AKT, float *zeta, uint32_t *flag, hipStream_t stream);

// Duplicate-boundary comparison of two records (wc_roms.hip): *key (set to
// all ones beforehand) receives min(point*3 + class) over the conflicts.
hipError_t launch_compare_records(const RecordPair<float> &F, int s_rho, size_t plane, unsigned long long *key,
                                  hipStream_t stream);

// One captured ensemble step: pack the record pair of steps[*stepIdx], move
// the population (block threads, ppt particles per thread, occupancy hint occ
// = 0 (none), 3 or 4 waves per SIMD), clear and fill the per-scenario surface counts
// (when surface != nullptr), advance *stepIdx.  Every kernel reads the step
// from device memory, so the same launch sequence is valid for any interval.
struct EnsLaunch {
    const EnsStep *steps;
    int *stepIdx;
    void *packBuf;
    GridPtrs G;
    EnsSoA s;
    int n;
    const float *sigma;
    int nScen;
    uint32_t *surface;
    unsigned long long *stepCounter;
    int s_w, s_rho, eta, xi;
    int block, ppt, occ;
    // cell-tile sort before the move every sortPeriod-th step (sortTile = 0:
    // never); tmp = scratch population, bins = (tiles + 1) counters
    int sortTile, sortPeriod;
    EnsSoA tmp;
    uint32_t *bins;
    // forcing staged in LDS per workgroup (move_ens_lds_kernel): chunk =
    // substeps between restagings (0: the plain kernel), halo in cells
    int ldsChunk, ldsHalo;
};
// LDS of one move_ens_lds_kernel workgroup (256 work-items): three per CU at
// occupancy hint 3, two otherwise (160 KB per CU, at most 64 KB per workgroup
// including its static LDS)
constexpr int kEnsLdsBytes3 = 47 * 1024;
constexpr int kEnsLdsBytes2 = 63 * 1024;
hipError_t launch_ens_step(int store, const EnsLaunch &L, hipStream_t stream);
// One release point of a device-generated ensemble population.
struct EnsSource {
    double k, j, i, spread, time;
    int scenario, subScenario;
    uint32_t subEvery;
};
hipError_t launch_ens_release(const EnsSource *src, const unsigned long long *first, int nsrc, long long n, int eta,
                              int xi, uint64_t seed, const EnsSoA &st, hipStream_t stream);

// ---- reconstruction network (wc_unet.hip) -----------------------------------
// 3x3 convolution as an implicit GEMM on BF16 MFMA.  Inputs x1 (c1 channels)
// and x2 (c2, may be 0) are NHWC, concatenated along channels, BF16 (or FP32
// when fp32_input, converted on load); m1 (optional, BF16, shaped as x1)
// masks x1 where m1 <= 0.  w_ = [nout][3][3][c1+c2] BF16, nout = cy1 + cy2.
// mode 0: y1 = BF16 relu(conv + bias) (cy2 = 0); mode 1: FP32 output
// channels [0, cy1) to y1 and [cy1, nout) to y2, each stored or accumulated.
struct ConvArgs {
    const void *x1;
    const void *x2;
    const uint16_t *m1;
    int c1, c2;
    const uint16_t *w_;
    const float *bias;
    void *y1;
    void *y2;
    int cy1, cy2;
    bool acc1, acc2;
    int mode;
    int n, h, w;
    // input gradient (mode 1) of a single-input layer, stored: with mo set
    // the result goes out as the previous layer's ReLU-masked gradient
    // mo[e] = bf16(my[e] > 0 ? g[e] : 0) (my: that layer's output); y1 is
    // still needed as FP32 scratch where the general kernel runs
    uint16_t *mo;
    const uint16_t *my;
    // with mo: this layer's input was the 2x2 average pool of the previous
    // layer's output, whose gradient is pg (FP32, twice the height and width)
    // plus a quarter of this one in each pooled pixel; mo, my and pg are at
    // that resolution (pg is read and, by the general kernel's second launch,
    // also updated)
    const float *pg;
    // input gradient (mode 1) of a two-input layer whose second input was the
    // 2x nearest upsampling of the previous layer's output: with uo set, that
    // part goes out as the previous layer's ReLU-masked gradient uo = bf16(uy >
    // 0 ? sum of the four pixels it fed : 0), at half the height and width; y2
    // (stored) and ud (FP32, half size) are scratch where the general kernel
    // or an odd row count runs
    uint16_t *uo;
    const uint16_t *uy;
    float *ud;
};
// A layer's ReLU-masked output gradient in BF16, written by the kernel that
// produces the gradient: gm[e] = bf16(y[e] > 0 ? g[e] : 0) in place of the
// FP32 gradient (gm null: the FP32 gradient is written)
struct MaskOut {
    uint16_t *gm;
    const uint16_t *y;
};
// dW[co][tap][ci] = sum_p dY[p][co] * (Y[p][co] > 0) * X[p + tap][ci] (dW and db
// are written, not accumulated)
struct WgradArgs {
    const float *dy;
    const uint16_t *y;
    const uint16_t *x1;
    const uint16_t *x2;
    int c1, c2, cout;
    float *dw;
    int n, h, w;
    int splits;
    // workspace of the tiled path (per-workgroup partial sums, then a
    // reduction): partCap floats; nullptr selects the atomic path
    float *part;
    size_t partCap;
    // the masked output gradient already in BF16 (or nullptr: dy and y)
    const uint16_t *dyb;
    // the bias gradient (optional, with dyb): db[co] = sum of dyb over the
    // pixels.  The tiled path sums it beside dW and needs db == dw + cout 9
    // cin (the parameter layout); the split-K path sums the partials
    // bpart[bnb][cout] of the masked gradient.
    float *db;
    const float *bpart;
    int bnb;
};
// Tiled weight gradients of several layers in one launch, then their slice
// sums in one more: each job has its own partial slices (a.part) and its
// workgroup range [first, next job's first); wgrad_plan fills a job.
constexpr int kWgradBatchMax = 6;
struct WgradJob {
    WgradArgs a;
    int R, rowsPerImg, stages, zn, gx, gy;
    int first;
};
struct WgradJobs {
    int nj;
    WgradJob job[kWgradBatchMax];
};
struct SliceSumJob {
    const float4 *part;
    float4 *dw;
    size_t E4;
    int zn, zg, first;
};
struct SliceSumJobs {
    int n;
    SliceSumJob s[kWgradBatchMax];
};
// The tiling of layer a (whose a.part is not yet set): false when the tiled
// path does not apply; *partFloats receives the partial floats it needs.
bool wgrad_plan(const WgradArgs &a, WgradJob &j, size_t *partFloats);
// b.job[0 .. b.nj) planned, each a.part set: the weight and bias gradients
// into each job's dW and db (written, not accumulated)
hipError_t launch_wgrad_batch(WgradJobs &b, hipStream_t s);
// floats of workspace that launch_wgrad's tiled path may use
constexpr size_t kWgradPartFloats = (size_t)1024 * 9 * 1024;
// A sum over workgroups that an earlier launch left as partials, finished by
// one workgroup of a later launch in a fixed order (so the result does not
// depend on scheduling): out[c] = scale * sum_{b < nb} part[b * stride + c]
// for c < m (m <= 256).  part null: no job.
struct PartialSums {
    const float *part;
    int nb, m, stride;
    float *out;
    float scale;
};
struct SideSums {
    PartialSums job[2];
};
hipError_t launch_conv3x3(const ConvArgs &a, bool fp32_input, hipStream_t s);
hipError_t launch_wgrad(const WgradArgs &a, hipStream_t s);
// whether launch_wgrad takes the tiled path (which sums the bias itself)
bool wgrad_is_tiled(const WgradArgs &a);
// out[e] = bf16(y[e] > 0 ? dy[e] : 0), a layer's ReLU-masked output gradient
// (E = pixels x cout, cout <= 256; dy null: out already holds it), and, when
// part is set, per workgroup the sums of that BF16 gradient per channel in
// part[b][cout] (partCap floats; *nblocks receives the number of
// workgroups).  The last workgroups also finish the jobs in `side`.
hipError_t launch_mask_bf16(const float *dy, const uint16_t *y, long long E, int cout, uint16_t *out, float *part,
                            size_t partCap, int *nblocks, const SideSums &side, hipStream_t s);
hipError_t launch_pool_fwd(const uint16_t *x, int n, int h, int w, int c, uint16_t *y, hipStream_t s);
hipError_t launch_pool_bwd(const float *dy, int n, int h, int w, int c, float *dx, bool accumulate, const MaskOut &m,
                           hipStream_t s);
hipError_t launch_up_fwd(const uint16_t *x, int n, int h, int w, int c, uint16_t *y, hipStream_t s);
hipError_t launch_up_bwd(const float *dy, int n, int h, int w, int c, float *dx, const MaskOut &m, hipStream_t s);
// x [n][cin][h][w] -> the first layer's NHWC BF16 input with c channels
// (P = n h w, hw = h w); x0 (optional) receives channel 0 as [n][h][w]
hipError_t launch_unet_input(const float *x, long long P, int hw, int cin, int c, uint16_t *y, float *x0,
                             hipStream_t s);
hipError_t launch_head_fwd(const uint16_t *d, const float *x, const float *wh, long long P, float *out,
                           hipStream_t s);
// The head's backward for loss = mean wgt (out - t)^2: dd (the gradient of
// the head's input d), or with gm set that gradient ReLU-masked by d in BF16
// (the last layer's masked gradient) in place of dd, and per workgroup the
// partials part[b][kHead + 2] of
// the head weights' and bias's gradients and of the loss (*nblocks
// workgroups), left for a later PartialSums job.
constexpr int kHeadPartials = 34;
hipError_t launch_head_bwd(const uint16_t *d, const float *out, const float *t, const float *wgt, const float *wh,
                           long long P, float *dd, uint16_t *gm, float *part, size_t partCap, int *nblocks,
                           hipStream_t s);
// the jobs of `side` (up to two), one workgroup each
hipError_t launch_finish_sums(const SideSums &side, hipStream_t s);
hipError_t launch_flip(const uint16_t *w, int cout, int cin, uint16_t *wf, hipStream_t s);
// Where each conv layer's weights sit: FP32 parameters [w_off, w_off +
// cout*9*cin) as [cout][9][cin]; their BF16 copy at wb_off of wb, and the
// same weights flipped and transposed for the data gradient,
// w'[ci][8 - tap][co], at wb_off of wf.
constexpr int kMaxConvLayers = 16;
struct WeightMap {
    int layers;
    int cout[kMaxConvLayers], cin[kMaxConvLayers];
    long long w_off[kMaxConvLayers], wb_off[kMaxConvLayers];
};
// Adam on every parameter, then the BF16 and flipped BF16 copies of the conv
// weights from the updated values (one pass).
hipError_t launch_adam_bf16(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                            float eps, int t, const WeightMap &map, uint16_t *wb, uint16_t *wf, hipStream_t s);
// The BF16 and flipped BF16 copies of the conv weights alone.
hipError_t launch_weights_bf16(const float *p, const WeightMap &map, uint16_t *wb, uint16_t *wf, hipStream_t s);
hipError_t launch_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                       float eps, int t, hipStream_t s);
hipError_t launch_to_bf16(const float *x, long long n, uint16_t *y, hipStream_t s);
hipError_t launch_fill_f32(float *a, long long n, float v, hipStream_t s);

}  // namespace wc
