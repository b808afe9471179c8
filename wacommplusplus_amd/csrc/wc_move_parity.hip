// FP64 parity move kernel for gfx950.
//
// One work-item per particle advances it over a whole forcing interval
// (cuda/kernel.cu:153-368 has the same granularity); the state stays in
// registers for all substeps.  The arithmetic follows the reference CPU path
// Particle::move (Particle.cpp:71-824) operation by operation, including the
// float casts of the time-interpolated samples and of the bilinear partial
// products, so that positions agree with the CPU to FP64 round-off:
//   - no contraction (file pragma + -ffp-contract=off), IEEE FP64 and FP32;
//   - the per-cell haversine sizes come from a table the host computes with
//     the same expression (Particle.cpp:544-574), or are recomputed here when
//     the caller passes only lon/lat (drop-in entry point);
//   - the keyed random draws reuse the SplitMix64 chain prefix
//     mix(mix(mix(seed)^id)^interval), which is the same for every substep of
//     the interval (NumericalHelpers.hpp:104-112), so a substep costs 7 mixes
//     instead of 30 and produces identical bits.
// Passive particles only (driftModel==0); the host rejects the rest.
#include "wc_device.h"
#include "wc_launch.h"

#pragma clang fp contract(off)

namespace wc {
namespace {

// packed gather layout of one interval (see PackedSampler)
struct alignas(16) PackUV {
    float v[4];      // u(t0), u(t1), v(t0), v(t1) at rho level k
};
struct alignas(32) PackWA {
    float v[8];      // w_k(t0), w_k(t1), w_k-1(t0), w_k-1(t1), akt_k(t0), akt_k(t1), akt_k-1(t0), akt_k-1(t1)
};
struct alignas(16) PackZD {
    float z0, z1;    // zeta(t0), zeta(t1)
    double h;
};

__device__ __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double std_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double sign_of(double a, double b) {
    return fabs(a) * (double)((b > 0) - (b < 0));
}
// validInterpolationCell on the floored coordinate, compared in double so a
// NaN or out-of-range coordinate is invalid (NumericalHelpers.hpp:12-14).
__device__ __forceinline__ bool valid_cell(double jf, double if_, int eta, int xi) {
    return jf >= 0.0 && if_ >= 0.0 && jf + 1.0 < (double)eta && if_ + 1.0 < (double)xi;
}
__device__ __forceinline__ double reflect_cell(double oldC, int oldCell, double cand, int candCell) {
    // NumericalHelpers.hpp:62-65
    if (candCell < oldCell) cand = oldCell + fabs(oldC - cand);
    else if (candCell > oldCell) cand = candCell - fmod(cand, 1.0);
    return cand;
}
__device__ __forceinline__ double reflect_domain(double c, double m) {
    // NumericalHelpers.hpp:67-74
    if (m <= 0) return 0;
    double r = fmod(c, 2.0 * m);
    if (r < 0) r += 2.0 * m;
    if (r >= m) r = 2.0 * m - r;
    if (r >= m) r = nextafter(m, 0.0);
    return r;
}

// (float)interpolateTime(a, b, alpha)  (Particle.cpp:186-202)
__device__ __forceinline__ float tl(float a, float b, double al) {
    return (float)((1.0 - al) * (double)a + al * (double)b);
}

// Corner values of a bilinear stencil in the reference's order
// (j,i), (j+1,i), (j+1,i+1), (j,i+1)  (Particle.cpp:271-274).
struct C4f {
    float a, b, c, d;
};
struct C4d {
    double a, b, c, d;
};

// float partial products summed in float, exactly as Particle.cpp:271-277
__device__ __forceinline__ float bil(const C4f &q, double iF, double jF) {
    float t1 = q.a * (1.0 - iF) * (1.0 - jF);
    float t2 = q.b * (1.0 - iF) * jF;
    float t3 = q.c * iF * jF;
    float t4 = q.d * iF * (1.0 - jF);
    return t1 + t2 + t3 + t4;
}
// trilinear over two W levels (Particle.cpp:464-474, 490-500)
__device__ __forceinline__ float tril(const C4f &t, const C4f &b, double iF, double jF, double kF) {
    float t1 = t.a * (1.0 - iF) * (1.0 - jF) * (1.0 - kF);
    float t2 = t.b * (1.0 - iF) * jF * (1.0 - kF);
    float t3 = t.c * iF * jF * (1.0 - kF);
    float t4 = t.d * iF * (1.0 - jF) * (1.0 - kF);
    float t5 = b.a * (1.0 - iF) * (1.0 - jF) * kF;
    float t6 = b.b * (1.0 - iF) * jF * kF;
    float t7 = b.c * iF * jF * kF;
    float t8 = b.d * iF * (1.0 - jF) * kF;
    return t1 + t2 + t3 + t4 + t5 + t6 + t7 + t8;
}

// Sampler over the reference's own array layout (drop-in entry point):
// index4 = ((t*levels + k+levels-1)*eta + j)*xi + i per record.  The cell
// metric comes from a {dx, dy} table the host computed with the expression
// of Particle.cpp:544-574 (as for the context path).
struct RawSampler {
    RecordPair<float> F;
    GridPtrs G;
    const double2 *md;
    int xi;
    size_t plane;
    int s_rho, s_w;

    __device__ __forceinline__ size_t lvl(int k, int levels) const { return (size_t)(k + levels - 1) * plane; }
    __device__ __forceinline__ C4f at4(const float *f0, const float *f1, size_t base, int j, int i, double al) const {
        const size_t q = base + (size_t)j * xi + i;
        return C4f{tl(f0[q], f1[q], al), tl(f0[q + xi], f1[q + xi], al), tl(f0[q + xi + 1], f1[q + xi + 1], al),
                   tl(f0[q + 1], f1[q + 1], al)};
    }
    __device__ __forceinline__ void zh4(int j, int i, double al, C4f &z, C4d &h) const {
        z = at4(F.zeta0, F.zeta1, 0, j, i, al);
        const size_t q = (size_t)j * xi + i;
        h = C4d{G.h[q], G.h[q + xi], G.h[q + xi + 1], G.h[q + 1]};
    }
    __device__ __forceinline__ void uvw4(int kI, int j, int i, double al, C4f &u, C4f &v, C4f &wt, C4f &wb) const {
        u = at4(F.u0, F.u1, lvl(kI, s_rho), j, i, al);
        v = at4(F.v0, F.v1, lvl(kI, s_rho), j, i, al);
        wt = at4(F.w0, F.w1, lvl(kI, s_w), j, i, al);
        wb = at4(F.w0, F.w1, lvl(kI - 1, s_w), j, i, al);
    }
    __device__ __forceinline__ void akt4(int kI, int j, int i, double al, C4f &at, C4f &ab) const {
        at = at4(F.akt0, F.akt1, lvl(kI, s_w), j, i, al);
        ab = at4(F.akt0, F.akt1, lvl(kI - 1, s_w), j, i, al);
    }
    __device__ __forceinline__ double depth(int kI) const { return G.depth[kI + s_w - 2]; }
    __device__ __forceinline__ void metric(int j, int i, double &dx, double &dy) const {
        const double2 m = md[(size_t)j * xi + i];
        dx = m.x;
        dy = m.y;
    }
};

// Content hash of a lon/lat grid: sum over cells of a SplitMix64 mix of each
// value's bits and its index, added once per wave into *hash (zeroed before).
__global__ void grid_hash_kernel(const double *__restrict__ lon, const double *__restrict__ lat, size_t n,
                                 unsigned long long *hash) {
    unsigned long long h = 0;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        h += mix64((uint64_t)__double_as_longlong(lon[q]) ^ (q * 0x9E3779B97F4A7C15ULL));
        h += mix64((uint64_t)__double_as_longlong(lat[q]) ^ ((q + n) * 0x9E3779B97F4A7C15ULL));
    }
    wave_add_u64(hash, h);
}

// Sampler over the context's packed interval layout: one 16-byte load gives
// u and v of both records at a corner, one 32-byte load w and AKt of both
// records at both W levels, one 16-byte load zeta of both records and h.  The
// values are the same FP32/FP64 numbers as the raw layout, so results are
// bit-identical; only the number of load instructions changes.
struct PackedSampler {
    const PackUV *uv;
    const PackWA *wa;
    const PackZD *z;
    const double2 *md;     // {dx, dy} haversine table
    const double *depthTab;
    int xi, nlev, s_w;
    size_t plane;

    __device__ __forceinline__ void zh4(int j, int i, double al, C4f &zz, C4d &h) const {
        const size_t q = (size_t)j * xi + i;
        const PackZD A = z[q], B = z[q + xi], C = z[q + xi + 1], D = z[q + 1];
        zz = C4f{tl(A.z0, A.z1, al), tl(B.z0, B.z1, al), tl(C.z0, C.z1, al), tl(D.z0, D.z1, al)};
        h = C4d{A.h, B.h, C.h, D.h};
    }
    __device__ __forceinline__ void uvw4(int kI, int j, int i, double al, C4f &u, C4f &v, C4f &wt, C4f &wb) const {
        const size_t q = (size_t)(kI + nlev - 1) * plane + (size_t)j * xi + i;
        const PackUV A = uv[q], B = uv[q + xi], C = uv[q + xi + 1], D = uv[q + 1];
        u = C4f{tl(A.v[0], A.v[1], al), tl(B.v[0], B.v[1], al), tl(C.v[0], C.v[1], al), tl(D.v[0], D.v[1], al)};
        v = C4f{tl(A.v[2], A.v[3], al), tl(B.v[2], B.v[3], al), tl(C.v[2], C.v[3], al), tl(D.v[2], D.v[3], al)};
        const PackWA E = wa[q], G2 = wa[q + xi], H = wa[q + xi + 1], K = wa[q + 1];
        wt = C4f{tl(E.v[0], E.v[1], al), tl(G2.v[0], G2.v[1], al), tl(H.v[0], H.v[1], al), tl(K.v[0], K.v[1], al)};
        wb = C4f{tl(E.v[2], E.v[3], al), tl(G2.v[2], G2.v[3], al), tl(H.v[2], H.v[3], al), tl(K.v[2], K.v[3], al)};
    }
    __device__ __forceinline__ void akt4(int kI, int j, int i, double al, C4f &at, C4f &ab) const {
        const size_t q = (size_t)(kI + nlev - 1) * plane + (size_t)j * xi + i;
        const PackWA E = wa[q], G2 = wa[q + xi], H = wa[q + xi + 1], K = wa[q + 1];
        at = C4f{tl(E.v[4], E.v[5], al), tl(G2.v[4], G2.v[5], al), tl(H.v[4], H.v[5], al), tl(K.v[4], K.v[5], al)};
        ab = C4f{tl(E.v[6], E.v[7], al), tl(G2.v[6], G2.v[7], al), tl(H.v[6], H.v[7], al), tl(K.v[6], K.v[7], al)};
    }
    __device__ __forceinline__ double depth(int kI) const { return depthTab[kI + s_w - 2]; }
    __device__ __forceinline__ void metric(int j, int i, double &dx, double &dy) const {
        const double2 m = md[(size_t)j * xi + i];
        dx = m.x;
        dy = m.y;
    }
};

// Box-Muller of two keyed uniforms (NumericalHelpers.hpp:114-120) from the
// substep key; component c uses uniforms 2c and 2c+1.
__device__ __forceinline__ double keyed_normal(uint64_t substepKey, uint64_t c) {
    double u1 = (double)(mix64(substepKey ^ (2 * c)) >> 11) * 0x1.0p-53;
    u1 = std_max(u1, 2.2250738585072014e-308);
    double u2 = (double)(mix64(substepKey ^ (2 * c + 1)) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.28318530717958647692 * u2);
}

// advectiveVelocity, passive branch (Particle.cpp:352-475)
template <class S>
__device__ __forceinline__ void velocity(const S &smp, int kI, double kF, int jI, double jF, int iI, double iF,
                                         double a, float &uu, float &vv, float &ww) {
    C4f u, v, wt, wb;
    smp.uvw4(kI, jI, iI, a, u, v, wt, wb);
    uu = bil(u, iF, jF);
    vv = bil(v, iF, jF);
    ww = tril(wt, wb, iF, jF, kF);
}

// Stand-in for a health value exp(-age/tau0) not yet computed (see ageLazy).
constexpr double kHealthPending = 2.0;

// Particle::move for one particle (Particle.cpp:71-824); returns the number of
// live substeps taken.  With Ev, the closures that act are OR-ed into ev
// (WC_EVENT_* bits); they change nothing else.
template <class Smp, bool Ev = false>
__device__ unsigned move_one(const MoveParams &P, const Smp &S, State &p, uint32_t &ev) {
    unsigned steps = 0;
    const double L = P.intervalLength;
    double elapsed = P.elapsed0;
    const double kLL = P.kLowerLimit;
    const int eta = P.eta, xi = P.xi;

    if (p.k > 0) {                                          // Particle.cpp:134-153
        if (Ev) ev |= WC_EVENT_VERTICAL;
        if (P.upper == WC_CLOSURE_CONSTRAINT) p.k = 0;
        else if (P.upper == WC_CLOSURE_KILL) p.health = -1;
        else if (P.upper == WC_CLOSURE_REFLECTION) p.k = -p.k;
    }
    if (p.k < kLL) {                                        // :156-176
        if (Ev) ev |= WC_EVENT_VERTICAL;
        if (P.lower == WC_CLOSURE_CONSTRAINT) p.k = kLL;
        else if (P.lower == WC_CLOSURE_KILL) p.health = -1;
        else if (P.lower == WC_CLOSURE_REFLECTION) p.k = 2.0 * kLL - p.k;
    }
    const double direction = P.direction;
    bool pending = false;                                   // p.health holds kHealthPending
    // random-key prefix, constant over the interval
    const uint64_t keyPrefix = mix64(mix64(mix64(P.seed) ^ p.id) ^ (uint64_t)P.intervalKey);

    while (elapsed < L) {                                   // :178
        const double stepDt = std_min(P.dti, L - elapsed);
        const double physicalTime = P.intervalStart + P.dirInt * elapsed;
        const double alpha = L == 0 ? 0 : (elapsed + 0.5 * stepDt) / L;
        const double alphaStart = L == 0 ? 0 : elapsed / L;

        if (P.dirInt > 0 && p.time > physicalTime) {        // :211-215
            elapsed = std_min(L, p.time - P.intervalStart);
            continue;
        }
        if (p.health < P.survprob) { p.health = -1; break; }   // :218-223

        const double kd = ceil(p.k);                        // :231-238
        const double kF = ceil(p.k) - p.k;
        const double jd = floor(p.j), id = floor(p.i);
        if (!valid_cell(jd, id, eta, xi)) { p.health = -1; break; }   // :241-248
        // level guard of cuda/kernel.cu:195 (only non-finite k reaches it)
        if (!(kd <= 0.0 && kd >= kLL)) { p.health = -1; break; }
        ++steps;
        const int kI = (int)kd, jI = (int)jd, iI = (int)id;
        const double jF = p.j - jI, iF = p.i - iI;

        // zeta, h at the particle (:271-290)
        C4f zc;
        C4d hcn;
        S.zh4(jI, iI, alpha, zc, hcn);
        const float zz = bil(zc, iF, jF);
        double h1 = hcn.a * (1.0 - iF) * (1.0 - jF);
        double h2 = hcn.b * (1.0 - iF) * jF;
        double h3 = hcn.c * iF * jF;
        double h4 = hcn.d * iF * (1.0 - jF);
        double hh = h1 + h2 + h3 + h4;
        double hc = hh + zz;
        if (hc <= P.shoreLimit && P.horizontal == WC_CLOSURE_KILL) p.health = -1;   // :311-338
        if (Ev && hc <= P.shoreLimit) ev |= WC_EVENT_SHALLOW;
        if (Ev && (hcn.a == 0 || hcn.b == 0 || hcn.c == 0 || hcn.d == 0)) ev |= WC_EVENT_COAST;

        float uu, vv, ww;                                   // :479-481
        velocity(S, kI, kF, jI, jF, iI, iF, P.midpoint ? alphaStart : alpha, uu, vv, ww);

        double dxleap = direction * uu * stepDt;            // :507-510
        double dyleap = direction * vv * stepDt;
        double dzleap = direction * (P.sv + ww) * stepDt;
        double rxleap = 0, ryleap = 0, rzleap = 0;
        if (P.randomOn) {                                   // :490-500, 518-533
            C4f at, ab;
            S.akt4(kI, jI, iI, alpha, at, ab);
            const float aa = tril(at, ab, iF, jF, kF);
            const double sigmaDepth = P.sigma * (1 - p.k / kLL);
            // floor(elapsed / dti): where elapsed is exactly q dti (the FMA
            // residual is 0), the quotient is exactly q; otherwise divide
            double q = rint(elapsed * P.invDti);
            if (fma(q, P.dti, -elapsed) != 0.0) q = floor(elapsed / P.dti);
            const uint64_t substep = (uint64_t)q;
            // sqrt(stepDt / dti) is exactly 1 for a whole step
            const double stochasticScale = stepDt == P.dti ? 1.0 : sqrt(stepDt / P.dti);
            const uint64_t sk = mix64(keyPrefix ^ substep);
            rxleap = sigmaDepth * keyed_normal(sk, 0) * stochasticScale;
            ryleap = sigmaDepth * keyed_normal(sk, 1) * stochasticScale;
            rzleap = sigmaDepth * keyed_normal(sk, 2) * aa * P.crid * stochasticScale;
        }
        double xleap = dxleap + rxleap;
        double yleap = dyleap + ryleap;
        double zleap = dzleap + rzleap;

        double xdist, ydist;                                // :544-579
        S.metric(jI, iI, xdist, ydist);
        const double zdist = hc * S.depth(kI);

        if (P.midpoint) {                                   // :585-641
            double kHalfLeap = direction * (P.sv + ww) * 0.5 * stepDt;
            if (fabs(kHalfLeap) > zdist) kHalfLeap = sign_of(zdist, kHalfLeap);
            const double iMid = p.i + direction * uu * 0.5 * stepDt / xdist;
            const double jMid = p.j + direction * vv * 0.5 * stepDt / ydist;
            const double kMid = std_min(0.0, std_max(kLL, p.k + kHalfLeap / zdist));
            const double jMd = floor(jMid), iMd = floor(iMid);
            bool midValid = valid_cell(jMd, iMd, eta, xi);
            double hcMid = 0;
            int iMI = 0, jMI = 0;
            double iMF = 0, jMF = 0;
            if (midValid) {
                iMI = (int)iMd; jMI = (int)jMd;
                iMF = iMid - iMI; jMF = jMid - jMI;
                C4f zm;
                C4d hm;
                S.zh4(jMI, iMI, alpha, zm, hm);
                hcMid = hm.a * (1.0 - iMF) * (1.0 - jMF) + hm.b * (1.0 - iMF) * jMF +
                        hm.c * iMF * jMF + hm.d * iMF * (1.0 - jMF) +
                        zm.a * (1.0 - iMF) * (1.0 - jMF) + zm.b * (1.0 - iMF) * jMF +
                        zm.c * iMF * jMF + zm.d * iMF * (1.0 - jMF);
                midValid = hcMid > P.shoreLimit && hcMid > 0;
            }
            if (Ev && !midValid) ev |= WC_EVENT_FALLBACK;
            // second stage at the midpoint, or the Euler fallback at the
            // start position; both sample at alpha, so one gather serves
            // either (the lanes of a wave may take different ones)
            const int kS = midValid ? (int)ceil(kMid) : kI;
            const double kSF = midValid ? ceil(kMid) - kMid : kF;
            const int jS = midValid ? jMI : jI, iS = midValid ? iMI : iI;
            const double jSF = midValid ? jMF : jF, iSF = midValid ? iMF : iF;
            float us, vs, ws;
            velocity(S, kS, kSF, jS, jSF, iS, iSF, alpha, us, vs, ws);
            if (midValid) {
                double xdistMid, ydistMid;
                S.metric(jMI, iMI, xdistMid, ydistMid);
                const double zdistMid = hcMid * S.depth(kS);
                dxleap = direction * us * stepDt * xdist / xdistMid;
                dyleap = direction * vs * stepDt * ydist / ydistMid;
                dzleap = direction * (P.sv + ws) * stepDt * zdist / zdistMid;
            } else {
                dxleap = direction * us * stepDt;
                dyleap = direction * vs * stepDt;
                dzleap = direction * (P.sv + ws) * stepDt;
            }
            xleap = dxleap + rxleap;
            yleap = dyleap + ryleap;
            zleap = dzleap + rzleap;
        }

        if (fabs(zleap) > zdist) zleap = sign_of(zdist, zleap);   // :644-648
        double jdet = p.j + yleap / ydist;                  // :651-657
        double idet = p.i + xleap / xdist;
        double kdet = p.k + zleap / zdist;

        if (kdet > 0.) {                                    // :665-683
            if (Ev) ev |= WC_EVENT_VERTICAL;
            if (P.upper == WC_CLOSURE_CONSTRAINT) kdet = 0;
            else if (P.upper == WC_CLOSURE_KILL) p.health = -1;
            else if (P.upper == WC_CLOSURE_REFLECTION) kdet = -kdet;
        }
        if (kdet < kLL) {                                   // :686-704
            if (Ev) ev |= WC_EVENT_VERTICAL;
            if (P.lower == WC_CLOSURE_CONSTRAINT) kdet = kLL;
            else if (P.lower == WC_CLOSURE_KILL) p.health = -1;
            else if (P.lower == WC_CLOSURE_REFLECTION) kdet = 2.0 * kLL - kdet;
        }

        const double jdd = floor(jdet), idd = floor(idet);  // :709-713
        if (valid_cell(jdd, idd, eta, xi)) {
            const int jdetI = (int)jdd, idetI = (int)idd;
            const double idetF = idet - idetI, jdetF = jdet - jdetI;   // :716-739
            C4f zd;
            C4d hd;
            S.zh4(jdetI, idetI, alpha, zd, hd);
            const float zzdet = bil(zd, idetF, jdetF);
            h1 = hd.a * (1.0 - idetF) * (1.0 - jdetF);
            h2 = hd.b * (1.0 - idetF) * jdetF;
            h3 = hd.c * idetF * jdetF;
            h4 = hd.d * idetF * (1.0 - jdetF);
            const double hcdet = (h1 + h2 + h3 + h4) + zzdet;
            if (hcdet <= P.shoreLimit) {                    // :753-771
                if (Ev) ev |= WC_EVENT_SHORE;
                if (P.horizontal == WC_CLOSURE_CONSTRAINT) { idet = p.i; jdet = p.j; }
                else if (P.horizontal == WC_CLOSURE_KILL) p.health = -1;
                else if (P.horizontal == WC_CLOSURE_REFLECTION) {
                    idet = reflect_cell(p.i, iI, idet, idetI);
                    jdet = reflect_cell(p.j, jI, jdet, jdetI);
                }
            }
            p.i = idet; p.j = jdet; p.k = kdet;             // :774-776
        } else {                                            // :777-791
            if (Ev) ev |= WC_EVENT_DOMAIN;
            if (P.horizontal == WC_CLOSURE_CONSTRAINT) p.k = kdet;
            else if (P.horizontal == WC_CLOSURE_KILL) p.health = -1;
            else if (P.horizontal == WC_CLOSURE_REFLECTION) {
                p.i = reflect_domain(idet, xi - 1);
                p.j = reflect_domain(jdet, eta - 1);
                p.k = kdet;
            }
        }
        if (p.health > 0) {                                 // :797-804
            p.age = p.age + stepDt;
            // health = exp(-age / tau0).  Below ageLazy the value is surely
            // >= survprob and > 0, so the stand-in kHealthPending (2, above
            // any survprob that permits deferring) answers both survival
            // tests alike, and the value is computed once, after the loop
            if (p.age < P.ageLazy) {
                p.health = kHealthPending;
                pending = true;
            } else {
                p.health = 1.0 * exp(-p.age / P.tau0);
                pending = false;
            }
            if (P.jibe && tag_type(p.tag) != 0 && jibe_flips(P, keyPrefix, elapsed, stepDt))   // :805-813
                p.tag = flip_side(p.tag);
        }
        elapsed += stepDt;                                  // :821
    }
    // a kill after the last deferral overwrote the stand-in with -1
    if (pending && p.health == kHealthPending) p.health = 1.0 * exp(-p.age / P.tau0);
    return steps;
}

// perm (optional): work-item t moves particle perm[t] (a cell-tile order)
template <class Smp, class Acc, bool Ev = false>
__global__ __launch_bounds__(256) void move_parity_kernel(MoveParams P, Smp S, Acc acc, int n,
                                                          unsigned long long *stepCounter, uint8_t *events,
                                                          const uint32_t *__restrict__ perm) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned steps = 0;
    if (t < n) {
        const int idx = perm ? (int)perm[t] : t;
        State p;
        acc.load(idx, p);
        uint32_t ev = 0;
        steps = move_one<Smp, Ev>(P, S, p, ev);
        acc.store(idx, p);
        if (Ev && ev) events[idx] |= (uint8_t)ev;
    }
    if (stepCounter) wave_add_u64(stepCounter, steps);
}

// Packs the interval's two FP32 records (and FP64 h) into the gather layout.
__global__ void pack_parity_kernel(RecordPair<float> F, const double *__restrict__ h, PackUV *__restrict__ uv,
                                   PackWA *__restrict__ wa, PackZD *__restrict__ z, int s_rho, int s_w,
                                   size_t plane) {
    const int nlev = s_w - 1;
    const size_t total = plane * nlev;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
        const size_t lv = q / plane, c = q - lv * plane;
        const int k = (int)lv - (nlev - 1);
        const size_t r = (size_t)(k + s_rho - 1) * plane + c;
        const size_t wt = (size_t)(k + s_w - 1) * plane + c, wb = wt - plane;
        uv[q] = PackUV{{F.u0[r], F.u1[r], F.v0[r], F.v1[r]}};
        wa[q] = PackWA{{F.w0[wt], F.w1[wt], F.w0[wb], F.w1[wb], F.akt0[wt], F.akt1[wt], F.akt0[wb], F.akt1[wb]}};
        if (lv == 0) z[c] = PackZD{F.zeta0[c], F.zeta1[c], h[c]};
    }
}

}  // namespace

size_t parity_pack_bytes(int s_w, size_t plane) {
    return plane * (size_t)(s_w - 1) * (sizeof(PackUV) + sizeof(PackWA)) + plane * sizeof(PackZD);
}

hipError_t launch_move_parity_soa(const MoveParams &P, const RecordPair<float> &F, const GridPtrs &G, void *packBuf,
                                  const double2 *metric, const SoA &s, int n, unsigned long long *stepCounter,
                                  bool events, const uint32_t *perm, int stages, hipStream_t stream) {
    const size_t plane = (size_t)P.eta * P.xi;
    const int nlev = P.s_w - 1;
    char *base = (char *)packBuf;
    PackedSampler S;
    S.uv = (const PackUV *)base;
    base += sizeof(PackUV) * plane * nlev;
    S.wa = (const PackWA *)base;
    base += sizeof(PackWA) * plane * nlev;
    S.z = (const PackZD *)base;
    S.md = metric;
    S.depthTab = G.depth;
    S.xi = P.xi;
    S.nlev = nlev;
    S.s_w = P.s_w;
    S.plane = plane;
    if (stages & kStagePack) {
        size_t blocks = (plane * nlev + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        pack_parity_kernel<<<(int)blocks, 256, 0, stream>>>(F, G.h, (PackUV *)S.uv, (PackWA *)S.wa, (PackZD *)S.z,
                                                            P.s_rho, P.s_w, plane);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!(stages & kStageMove) || n <= 0) return hipSuccess;
    const int block = 256;
    if (events)
        move_parity_kernel<PackedSampler, SoAAccess, true><<<(n + block - 1) / block, block, 0, stream>>>(
            P, S, SoAAccess{s, P.jibe}, n, stepCounter, s.ev, perm);
    else
        move_parity_kernel<PackedSampler, SoAAccess><<<(n + block - 1) / block, block, 0, stream>>>(
            P, S, SoAAccess{s, P.jibe}, n, stepCounter, nullptr, perm);
    return hipGetLastError();
}

hipError_t launch_grid_hash(const double *lon, const double *lat, size_t n, unsigned long long *hash,
                            hipStream_t stream) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    grid_hash_kernel<<<(int)(blocks ? blocks : 1), 256, 0, stream>>>(lon, lat, n, hash);
    return hipGetLastError();
}

hipError_t launch_move_parity_aos(const MoveParams &P, const RecordPair<float> &F, const GridPtrs &G,
                                  const double2 *metric, void *packBuf, wc_particle_data *p, int n,
                                  unsigned long long *stepCounter, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    if (!packBuf) {
        RawSampler S{F, G, metric, P.xi, (size_t)P.eta * P.xi, P.s_rho, P.s_w};
        move_parity_kernel<RawSampler, AoSAccess><<<(n + block - 1) / block, block, 0, stream>>>(
            P, S, AoSAccess{p}, n, stepCounter, nullptr, nullptr);
        return hipGetLastError();
    }
    // the interval pack of the two records, then the move through it: the
    // same FP32 values, so the same results, with a third of the gathers
    const size_t plane = (size_t)P.eta * P.xi;
    const int nlev = P.s_w - 1;
    char *base = (char *)packBuf;
    PackedSampler S;
    S.uv = (const PackUV *)base;
    base += sizeof(PackUV) * plane * nlev;
    S.wa = (const PackWA *)base;
    base += sizeof(PackWA) * plane * nlev;
    S.z = (const PackZD *)base;
    S.md = metric;
    S.depthTab = G.depth;
    S.xi = P.xi;
    S.nlev = nlev;
    S.s_w = P.s_w;
    S.plane = plane;
    size_t blocks = (plane * nlev + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    pack_parity_kernel<<<(int)blocks, 256, 0, stream>>>(F, G.h, (PackUV *)S.uv, (PackWA *)S.wa, (PackZD *)S.z,
                                                        P.s_rho, P.s_w, plane);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    move_parity_kernel<PackedSampler, AoSAccess><<<(n + block - 1) / block, block, 0, stream>>>(
        P, S, AoSAccess{p}, n, stepCounter, nullptr, nullptr);
    return hipGetLastError();
}

}  // namespace wc
