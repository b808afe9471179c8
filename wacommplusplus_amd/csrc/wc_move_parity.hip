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

#pragma clang fp contract(off)

namespace wc {
namespace {

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

// (float)interpolateTime(a[idx], b[idx], alpha)  (Particle.cpp:186-202)
__device__ __forceinline__ float tl(const float *__restrict__ a, const float *__restrict__ b, size_t idx,
                                    double al) {
    return (float)((1.0 - al) * (double)a[idx] + al * (double)b[idx]);
}

struct Sampler {
    const RecordPair<float> &F;
    const GridPtrs &G;
    int xi;
    size_t plane;
    int s_rho, s_w;

    __device__ __forceinline__ size_t lvl(int k, int levels, int j, int i) const {
        return (size_t)(k + levels - 1) * plane + (size_t)j * xi + i;
    }
    __device__ __forceinline__ float zeta(int j, int i, double a) const {
        return tl(F.zeta0, F.zeta1, (size_t)j * xi + i, a);
    }
    __device__ __forceinline__ double h(int j, int i) const { return G.h[(size_t)j * xi + i]; }

    // advectiveVelocity, passive branch (Particle.cpp:352-475)
    __device__ __forceinline__ void velocity(int kI, double kF, int jI, double jF, int iI, double iF, double a,
                                             float &uu, float &vv, float &ww) const {
        const double wi0 = 1.0 - iF, wj0 = 1.0 - jF;
        float u1 = tl(F.u0, F.u1, lvl(kI, s_rho, jI, iI), a) * wi0 * wj0;
        float u2 = tl(F.u0, F.u1, lvl(kI, s_rho, jI + 1, iI), a) * wi0 * jF;
        float u3 = tl(F.u0, F.u1, lvl(kI, s_rho, jI + 1, iI + 1), a) * iF * jF;
        float u4 = tl(F.u0, F.u1, lvl(kI, s_rho, jI, iI + 1), a) * iF * wj0;
        uu = u1 + u2 + u3 + u4;
        float v1 = tl(F.v0, F.v1, lvl(kI, s_rho, jI, iI), a) * wi0 * wj0;
        float v2 = tl(F.v0, F.v1, lvl(kI, s_rho, jI + 1, iI), a) * wi0 * jF;
        float v3 = tl(F.v0, F.v1, lvl(kI, s_rho, jI + 1, iI + 1), a) * iF * jF;
        float v4 = tl(F.v0, F.v1, lvl(kI, s_rho, jI, iI + 1), a) * iF * wj0;
        vv = v1 + v2 + v3 + v4;
        const double wk0 = 1.0 - kF;
        float w1 = tl(F.w0, F.w1, lvl(kI, s_w, jI, iI), a) * wi0 * wj0 * wk0;
        float w2 = tl(F.w0, F.w1, lvl(kI, s_w, jI + 1, iI), a) * wi0 * jF * wk0;
        float w3 = tl(F.w0, F.w1, lvl(kI, s_w, jI + 1, iI + 1), a) * iF * jF * wk0;
        float w4 = tl(F.w0, F.w1, lvl(kI, s_w, jI, iI + 1), a) * iF * wj0 * wk0;
        float w5 = tl(F.w0, F.w1, lvl(kI - 1, s_w, jI, iI), a) * wi0 * wj0 * kF;
        float w6 = tl(F.w0, F.w1, lvl(kI - 1, s_w, jI + 1, iI), a) * wi0 * jF * kF;
        float w7 = tl(F.w0, F.w1, lvl(kI - 1, s_w, jI + 1, iI + 1), a) * iF * jF * kF;
        float w8 = tl(F.w0, F.w1, lvl(kI - 1, s_w, jI, iI + 1), a) * iF * wj0 * kF;
        ww = w1 + w2 + w3 + w4 + w5 + w6 + w7 + w8;
    }

    // AKt at the particle (Particle.cpp:490-500)
    __device__ __forceinline__ float akt(int kI, double kF, int jI, double jF, int iI, double iF, double a) const {
        const double wi0 = 1.0 - iF, wj0 = 1.0 - jF, wk0 = 1.0 - kF;
        float a1 = tl(F.akt0, F.akt1, lvl(kI, s_w, jI, iI), a) * wi0 * wj0 * wk0;
        float a2 = tl(F.akt0, F.akt1, lvl(kI, s_w, jI + 1, iI), a) * wi0 * jF * wk0;
        float a3 = tl(F.akt0, F.akt1, lvl(kI, s_w, jI + 1, iI + 1), a) * iF * jF * wk0;
        float a4 = tl(F.akt0, F.akt1, lvl(kI, s_w, jI, iI + 1), a) * iF * wj0 * wk0;
        float a5 = tl(F.akt0, F.akt1, lvl(kI - 1, s_w, jI, iI), a) * wi0 * wj0 * kF;
        float a6 = tl(F.akt0, F.akt1, lvl(kI - 1, s_w, jI + 1, iI), a) * wi0 * jF * kF;
        float a7 = tl(F.akt0, F.akt1, lvl(kI - 1, s_w, jI + 1, iI + 1), a) * iF * jF * kF;
        float a8 = tl(F.akt0, F.akt1, lvl(kI - 1, s_w, jI, iI + 1), a) * iF * wj0 * kF;
        return a1 + a2 + a3 + a4 + a5 + a6 + a7 + a8;
    }

    // Haversine cell sizes (Particle.cpp:548-574 / 616-621) from the table or
    // recomputed from lon/lat with the same expression.
    __device__ __forceinline__ void metric(int j, int i, double &dx, double &dy) const {
        const size_t c = (size_t)j * xi + i;
        if (G.dx) {
            dx = G.dx[c];
            dy = G.dy[c];
            return;
        }
        double dLat = G.latRad[c + xi] - G.latRad[c];
        double t = pow(sin(0.5 * dLat), 2);
        dy = 2.0 * atan2(pow(t, .5), pow(1.0 - t, .5)) * 6371000.0;
        double dLon = G.lonRad[c + 1] - G.lonRad[c];
        t = pow(sin(0.5 * dLon), 2) * cos(G.latRad[c + 1]) * cos(G.latRad[c]);
        dx = 2.0 * atan2(pow(t, .5), pow(1.0 - t, .5)) * 6371000.0;
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

// Particle::move for one particle (Particle.cpp:71-824); returns the number of
// live substeps taken.
__device__ unsigned move_one(const MoveParams &P, const Sampler &S, State &p) {
    unsigned steps = 0;
    const double L = P.intervalLength;
    double elapsed = P.elapsed0;
    const double kLL = P.kLowerLimit;
    const int eta = P.eta, xi = P.xi;

    if (p.k > 0) {                                          // Particle.cpp:134-153
        if (P.upper == WC_CLOSURE_CONSTRAINT) p.k = 0;
        else if (P.upper == WC_CLOSURE_KILL) p.health = -1;
        else if (P.upper == WC_CLOSURE_REFLECTION) p.k = -p.k;
    }
    if (p.k < kLL) {                                        // :156-176
        if (P.lower == WC_CLOSURE_CONSTRAINT) p.k = kLL;
        else if (P.lower == WC_CLOSURE_KILL) p.health = -1;
        else if (P.lower == WC_CLOSURE_REFLECTION) p.k = 2.0 * kLL - p.k;
    }
    const double direction = P.direction;
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
        float z1 = S.zeta(jI, iI, alpha) * (1.0 - iF) * (1.0 - jF);
        float z2 = S.zeta(jI + 1, iI, alpha) * (1.0 - iF) * jF;
        float z3 = S.zeta(jI + 1, iI + 1, alpha) * iF * jF;
        float z4 = S.zeta(jI, iI + 1, alpha) * iF * (1.0 - jF);
        float zz = z1 + z2 + z3 + z4;
        double h1 = S.h(jI, iI) * (1.0 - iF) * (1.0 - jF);
        double h2 = S.h(jI + 1, iI) * (1.0 - iF) * jF;
        double h3 = S.h(jI + 1, iI + 1) * iF * jF;
        double h4 = S.h(jI, iI + 1) * iF * (1.0 - jF);
        double hh = h1 + h2 + h3 + h4;
        double hc = hh + zz;
        if (hc <= P.shoreLimit && P.horizontal == WC_CLOSURE_KILL) p.health = -1;   // :311-338

        float uu, vv, ww;                                   // :479-481
        S.velocity(kI, kF, jI, jF, iI, iF, P.midpoint ? alphaStart : alpha, uu, vv, ww);

        double dxleap = direction * uu * stepDt;            // :507-510
        double dyleap = direction * vv * stepDt;
        double dzleap = direction * (P.sv + ww) * stepDt;
        double rxleap = 0, ryleap = 0, rzleap = 0;
        if (P.randomOn) {                                   // :518-533
            const float aa = S.akt(kI, kF, jI, jF, iI, iF, alpha);
            const double sigmaDepth = P.sigma * (1 - p.k / kLL);
            const uint64_t substep = (uint64_t)floor(elapsed / P.dti);
            const double stochasticScale = sqrt(stepDt / P.dti);
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
        const double zdist = hc * S.G.depth[kI + P.s_w - 2];

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
                hcMid = S.h(jMI, iMI) * (1.0 - iMF) * (1.0 - jMF) + S.h(jMI + 1, iMI) * (1.0 - iMF) * jMF +
                        S.h(jMI + 1, iMI + 1) * iMF * jMF + S.h(jMI, iMI + 1) * iMF * (1.0 - jMF) +
                        S.zeta(jMI, iMI, alpha) * (1.0 - iMF) * (1.0 - jMF) +
                        S.zeta(jMI + 1, iMI, alpha) * (1.0 - iMF) * jMF +
                        S.zeta(jMI + 1, iMI + 1, alpha) * iMF * jMF +
                        S.zeta(jMI, iMI + 1, alpha) * iMF * (1.0 - jMF);
                midValid = hcMid > P.shoreLimit && hcMid > 0;
            }
            if (midValid) {
                const int kMI = (int)ceil(kMid);
                const double kMF = ceil(kMid) - kMid;
                float um, vm, wm;
                S.velocity(kMI, kMF, jMI, jMF, iMI, iMF, alpha, um, vm, wm);
                double xdistMid, ydistMid;
                S.metric(jMI, iMI, xdistMid, ydistMid);
                const double zdistMid = hcMid * S.G.depth[kMI + P.s_w - 2];
                dxleap = direction * um * stepDt * xdist / xdistMid;
                dyleap = direction * vm * stepDt * ydist / ydistMid;
                dzleap = direction * (P.sv + wm) * stepDt * zdist / zdistMid;
            } else {
                float ue, ve, we;
                S.velocity(kI, kF, jI, jF, iI, iF, alpha, ue, ve, we);
                dxleap = direction * ue * stepDt;
                dyleap = direction * ve * stepDt;
                dzleap = direction * (P.sv + we) * stepDt;
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
            if (P.upper == WC_CLOSURE_CONSTRAINT) kdet = 0;
            else if (P.upper == WC_CLOSURE_KILL) p.health = -1;
            else if (P.upper == WC_CLOSURE_REFLECTION) kdet = -kdet;
        }
        if (kdet < kLL) {                                   // :686-704
            if (P.lower == WC_CLOSURE_CONSTRAINT) kdet = kLL;
            else if (P.lower == WC_CLOSURE_KILL) p.health = -1;
            else if (P.lower == WC_CLOSURE_REFLECTION) kdet = 2.0 * kLL - kdet;
        }

        const double jdd = floor(jdet), idd = floor(idet);  // :709-713
        if (valid_cell(jdd, idd, eta, xi)) {
            const int jdetI = (int)jdd, idetI = (int)idd;
            const double idetF = idet - idetI, jdetF = jdet - jdetI;   // :716-739
            z1 = S.zeta(jdetI, idetI, alpha) * (1.0 - idetF) * (1.0 - jdetF);
            z2 = S.zeta(jdetI + 1, idetI, alpha) * (1.0 - idetF) * jdetF;
            z3 = S.zeta(jdetI + 1, idetI + 1, alpha) * idetF * jdetF;
            z4 = S.zeta(jdetI, idetI + 1, alpha) * idetF * (1.0 - jdetF);
            float zzdet = z1 + z2 + z3 + z4;
            h1 = S.h(jdetI, idetI) * (1.0 - idetF) * (1.0 - jdetF);
            h2 = S.h(jdetI + 1, idetI) * (1.0 - idetF) * jdetF;
            h3 = S.h(jdetI + 1, idetI + 1) * idetF * jdetF;
            h4 = S.h(jdetI, idetI + 1) * idetF * (1.0 - jdetF);
            double hcdet = (h1 + h2 + h3 + h4) + zzdet;
            if (hcdet <= P.shoreLimit) {                    // :753-771
                if (P.horizontal == WC_CLOSURE_CONSTRAINT) { idet = p.i; jdet = p.j; }
                else if (P.horizontal == WC_CLOSURE_KILL) p.health = -1;
                else if (P.horizontal == WC_CLOSURE_REFLECTION) {
                    idet = reflect_cell(p.i, iI, idet, idetI);
                    jdet = reflect_cell(p.j, jI, jdet, jdetI);
                }
            }
            p.i = idet; p.j = jdet; p.k = kdet;             // :774-776
        } else {                                            // :777-791
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
            p.health = 1.0 * exp(-p.age / P.tau0);
        }
        elapsed += stepDt;                                  // :821
    }
    return steps;
}

template <class Acc>
__global__ __launch_bounds__(256) void move_parity_kernel(MoveParams P, RecordPair<float> F, GridPtrs G, Acc acc,
                                                          int n, unsigned long long *stepCounter) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned steps = 0;
    if (idx < n) {
        Sampler S{F, G, P.xi, (size_t)P.eta * P.xi, P.s_rho, P.s_w};
        State p;
        acc.load(idx, p);
        steps = move_one(P, S, p);
        acc.store(idx, p);
    }
    if (stepCounter) wave_add_u64(stepCounter, steps);
}

}  // namespace

hipError_t launch_move_parity_soa(const MoveParams &P, const RecordPair<float> &F, const GridPtrs &G, const SoA &s,
                                  int n, unsigned long long *stepCounter, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int grid = (n + block - 1) / block;
    move_parity_kernel<SoAAccess><<<grid, block, 0, stream>>>(P, F, G, SoAAccess{s}, n, stepCounter);
    return hipGetLastError();
}

hipError_t launch_move_parity_aos(const MoveParams &P, const RecordPair<float> &F, const GridPtrs &G,
                                  wc_particle_data *p, int n, unsigned long long *stepCounter, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int block = 256;
    const int grid = (n + block - 1) / block;
    move_parity_kernel<AoSAccess><<<grid, block, 0, stream>>>(P, F, G, AoSAccess{p}, n, stepCounter);
    return hipGetLastError();
}

}  // namespace wc
