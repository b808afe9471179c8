// ROMS record normalisation on the device (ROMSAdapter.cpp:56-300, 309-368,
// 448-466).
//
// One forcing record of a ROMS file arrives as u on the u grid
// [s_rho][eta_u][xi_u], v on the v grid [s_rho][eta_v][xi_v], and w, AKt
// [s_w][eta][xi], zeta [eta][xi] on the rho grid.  The kernel turns it into
// the rho-grid record the move reads, writing straight into a forcing slot:
//   - U, V: float mean of the wet faces beside each wet rho point (uv2rho,
//     ROMSAdapter.cpp:309-368), then rotated by the grid angle into
//     east/north in double and rounded to float (:276-300); 0 on land;
//   - W, AKT: zeroed on land in place (wakt2wakt, :448-466);
//   - zeta: zeroed on land or where >= 1e36 in place (:254-264);
//   - the checks the reference makes before it normalises (:152-166, 290-298):
//     a wet u or v face holding a non-finite or missing value, or a rotated
//     velocity outside float range, sets a bit of *flag.
// One work-item per rho point and 4 of its levels: it reads the point's
// mask byte and rotation once for the 4 levels, and issues all its loads
// before the first use (one memory round trip).  Per 1024x768x11 record:
// 55 us, against 63 us with one level per work-item (70 us with its loads
// not hoisted) and 58 us with all 12 levels per work-item (a loop over the
// column without hoisted loads ran at ~30 % of peak).  An XCD-aware mapping
// that swept the levels of a tile in one L2 measured slower (65 us).  The
// three masks are folded at set-up into one byte
// per rho point (kFace* bits: the point is wet, and which of its four faces
// are wet), so a work-item reads 1 B of mask state.  Consecutive work-items
// take consecutive i, so every load and store coalesces.  The u/v grids are
// no larger than the rho grid (ROMS staggering: eta_u = eta, xi_u = xi-1,
// eta_v = eta-1, xi_v = xi), so every face is checked by the work-item of
// the rho point with the same (j, i), exactly once.
// The work is HBM-bound byte traffic: each raw u/v value is read by at most
// two neighbouring work-items (the second read is a cache hit), U/V are
// written once, and W, AKT and zeta are only rewritten at land points.
#include <cfloat>

#include "wc_launch.h"

#pragma clang fp contract(off)

namespace wc {
namespace {

constexpr int kBlock = 256;
constexpr int kLevelsPerItem = 4;     // levels of one rho point per work-item

__device__ __forceinline__ bool valid_value(float value, const double *missing, int n) {
    const double x = value;
    if (!isfinite(x)) return false;
    for (int q = 0; q < n; ++q)
        if (x == missing[q]) return false;
    return true;
}

// uv2rho and the rotation for one rho point of one level (ROMSAdapter.cpp:
// 309-368, 276-300); the face values arrive already loaded.
__device__ __forceinline__ void normalize_uv(const RomsDims &d, const RomsMissing &miss, uint32_t f, double2 cs,
                                             float u0, float uw, float v0, float vs, float *U, float *V,
                                             uint32_t *flag) {
    const bool wet = f & kFaceWet;
    uint32_t bad = 0;
    if ((f & kFaceU) && !valid_value(u0, miss.u, miss.nu)) bad |= 1u;
    if ((f & kFaceV) && !valid_value(v0, miss.v, miss.nv)) bad |= 2u;
    float east_f = 0.0f, north_f = 0.0f;
    if (wet) {
        float usum = 0.0f, vsum = 0.0f;
        int ucount = 0, vcount = 0;
        if (f & kFaceU) { usum += u0; ucount++; }
        if (f & kFaceUWest) { usum += uw; ucount++; }
        if (f & kFaceV) { vsum += v0; vcount++; }
        if (f & kFaceVSouth) { vsum += vs; vcount++; }
        const float gu_f = ucount > 0 ? usum / ucount : 0.0f;
        const float gv_f = vcount > 0 ? vsum / vcount : 0.0f;
        const double gu = gu_f, gv = gv_f;
        double east = gu, north = gv;
        if (!d.earth) {
            east = gu * cs.x - gv * cs.y;
            north = gu * cs.y + gv * cs.x;
        }
        if (!isfinite(east) || !isfinite(north) || fabs(east) > (double)FLT_MAX || fabs(north) > (double)FLT_MAX) {
            bad |= 4u;
            east_f = gu_f;                              // the reference leaves the mean and throws
            north_f = gv_f;
        } else {
            east_f = (float)east;
            north_f = (float)north;
        }
    }
    *U = east_f;
    *V = north_f;
    if (bad) atomicOr(flag, bad);
}

__global__ __launch_bounds__(kBlock) void roms_normalize_kernel(RomsDims d, RomsMissing miss,
                                                                const uint8_t *__restrict__ face,
                                                                const double2 *__restrict__ rot,
                                                                const float *__restrict__ u,
                                                                const float *__restrict__ v, float *__restrict__ U,
                                                                float *__restrict__ V, float *__restrict__ W,
                                                                float *__restrict__ AKT, float *__restrict__ zeta,
                                                                uint32_t *__restrict__ flag) {
    const size_t plane = (size_t)d.eta * d.xi;
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= plane) return;
    const int k0 = blockIdx.y * kLevelsPerItem;    // storage levels k0 .. k0 + kLevelsPerItem - 1, 0 = bottom
    const int j = (int)(c / d.xi), i = (int)(c - (size_t)j * d.xi);

    // Every load is issued before the first use, guarded only by index
    // bounds (no memory dependence), so a work-item waits for one memory
    // round trip; the face bits select what the arithmetic uses.
    const bool in_u = j < d.eta_u && i < d.xi_u;
    const bool in_uw = j < d.eta_u && i > 0 && i - 1 < d.xi_u;
    const bool in_v = j < d.eta_v && i < d.xi_v;
    const bool in_vs = j > 0 && j - 1 < d.eta_v && i < d.xi_v;
    const uint32_t f = face[c];
    const double2 cs = (k0 < d.s_rho && !d.earth) ? rot[c] : make_double2(1.0, 0.0);   // {cos, sin}; 1, 0 on land
    float u0[kLevelsPerItem], uw[kLevelsPerItem], v0[kLevelsPerItem], vs[kLevelsPerItem];
#pragma unroll
    for (int q = 0; q < kLevelsPerItem; ++q) {
        const int k = k0 + q;
        const bool rho_level = k < d.s_rho;
        const float *uk = u + ((size_t)k * d.eta_u + j) * d.xi_u + i;     // face (j, i) of level k
        const float *vk = v + ((size_t)k * d.eta_v + j) * d.xi_v + i;
        u0[q] = rho_level && in_u ? uk[0] : 0.0f;
        uw[q] = rho_level && in_uw ? uk[-1] : 0.0f;
        v0[q] = rho_level && in_v ? vk[0] : 0.0f;
        vs[q] = rho_level && in_vs ? vk[-(ptrdiff_t)d.xi_v] : 0.0f;
    }
    const float z = k0 == 0 ? zeta[c] : 0.0f;
    const bool wet = f & kFaceWet;
#pragma unroll
    for (int q = 0; q < kLevelsPerItem; ++q) {
        const int k = k0 + q;
        if (k > d.s_rho) break;                 // W and AKT have s_rho + 1 levels
        const size_t o = (size_t)k * plane + c;
        if (!wet) {
            W[o] = 0.0f;
            AKT[o] = 0.0f;
        }
        if (k < d.s_rho) normalize_uv(d, miss, f, cs, u0[q], uw[q], v0[q], vs[q], U + o, V + o, flag);
    }
    if (k0 == 0 && (!wet || z >= 1e+36f)) zeta[c] = 0.0f;
}

// The duplicate-boundary check of OceanModelAdapter::appendBoundaryRecord
// (OceanModelAdapter.cpp:216-233): two records must agree under float `!=`
// (so a NaN always conflicts).  The reference loops over points (j, i) and,
// per point, checks zeta, then u/v over the levels, then w/AKt, throwing at
// the first conflict; the conflict it reports is therefore the one with the
// smallest key point*3 + class (class 0 zeta, 1 horizontal, 2 vertical).
// One work-item per (row, point), row 0 zeta, rows 1..s_rho u/v of a level,
// the rest w/AKt of a level; each wave offers its lowest conflicting lane
// (the smallest point) to one atomicMin.
__global__ __launch_bounds__(kBlock) void compare_records_kernel(RecordPair<float> F, int s_rho, size_t plane,
                                                                 unsigned long long *__restrict__ key) {
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const int row = blockIdx.y;
    bool conflict = false;
    unsigned cls = 0;
    if (c < plane) {
        if (row == 0) {
            conflict = F.zeta0[c] != F.zeta1[c];
        } else if (row <= s_rho) {
            const size_t o = (size_t)(row - 1) * plane + c;
            conflict = F.u0[o] != F.u1[o] || F.v0[o] != F.v1[o];
            cls = 1;
        } else {
            const size_t o = (size_t)(row - 1 - s_rho) * plane + c;
            conflict = F.w0[o] != F.w1[o] || F.akt0[o] != F.akt1[o];
            cls = 2;
        }
    }
    const unsigned long long lanes = __ballot(conflict);
    if (lanes && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)lanes) - 1))
        atomicMin(key, (unsigned long long)c * 3ull + cls);
}

}  // namespace

hipError_t launch_compare_records(const RecordPair<float> &F, int s_rho, size_t plane, unsigned long long *key,
                                  hipStream_t stream) {
    if (plane == 0) return hipSuccess;
    const dim3 grid((unsigned)((plane + kBlock - 1) / kBlock), (unsigned)(2 * s_rho + 2));
    compare_records_kernel<<<grid, kBlock, 0, stream>>>(F, s_rho, plane, key);
    return hipGetLastError();
}

hipError_t launch_roms_normalize(const RomsDims &d, const RomsMissing &miss, const uint8_t *face,
                                 const double2 *rot, const float *u, const float *v, float *U, float *V, float *W,
                                 float *AKT, float *zeta, uint32_t *flag, hipStream_t stream) {
    const size_t plane = (size_t)d.eta * d.xi;
    if (plane == 0) return hipSuccess;
    const dim3 grid((unsigned)((plane + kBlock - 1) / kBlock),
                    (unsigned)((d.s_rho + 1 + kLevelsPerItem - 1) / kLevelsPerItem));
    roms_normalize_kernel<<<grid, kBlock, 0, stream>>>(d, miss, face, rot, u, v, U, V, W, AKT, zeta, flag);
    return hipGetLastError();
}

}  // namespace wc
