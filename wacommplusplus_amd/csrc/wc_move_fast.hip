// Mixed-precision move kernel for gfx950: the same operator as Particle::move
// (Particle.cpp:71-824: time-lerped bilinear u, v at the containing rho level,
// trilinear w and AKt over the two W levels, explicit midpoint or Euler step,
// keyed random displacement, surface/bottom/shore/open-boundary closures,
// exponential decay), computed in FP32 from FP32, FP16 or BF16 forcing.
//
// State: each work-item keeps its particle as an integer cell (jI, iI) plus
// FP32 offsets in [0,1) and an FP32 level k, for a whole forcing interval in
// registers; only elapsed time and age stay FP64.  The FP64 SoA state is read
// once and written once per interval (j = jI + offset is exact in FP64).
// Health is evaluated once per interval: exp(-age/tau0) < survprob is
// equivalent to age > -tau0*log(survprob), tested every substep.
//
// Diffusion draws: Philox4x32-10 with key = seed, counter = (substep,
// interval key, id lo, id hi); one call gives the four uniforms of two
// Box-Muller pairs (three normals used).  Statistically equivalent to the
// reference's SplitMix64 draws, not draw-for-draw identical.
#include <hip/hip_fp16.h>

#include "wc_device.h"
#include "wc_philox.h"

#pragma clang fp contract(fast)

namespace wc {
namespace {

// storage -> float
template <class T>
__device__ __forceinline__ float ld(const T *__restrict__ p, size_t i);
template <>
__device__ __forceinline__ float ld<float>(const float *__restrict__ p, size_t i) {
    return p[i];
}
template <>
__device__ __forceinline__ float ld<_Float16>(const _Float16 *__restrict__ p, size_t i) {
    return (float)p[i];
}
template <>
__device__ __forceinline__ float ld<uint16_t>(const uint16_t *__restrict__ p, size_t i) {
    return __uint_as_float((uint32_t)p[i] << 16);   // bfloat16 bits
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

template <class T>
struct FastSampler {
    const T *zeta0, *zeta1, *u0, *u1, *v0, *v1, *w0, *w1, *a0, *a1;
    const float *hf, *idxf, *idyf, *depthf;
    int xi;
    size_t plane;
    int s_rho, s_w;

    // time-lerped bilinear sample of one level plane at base offset `b`
    // (index of corner (jI, iI)); fi, fj offsets in [0,1)
    __device__ __forceinline__ float bil(const T *__restrict__ f0, const T *__restrict__ f1, size_t b, float fi,
                                         float fj, float al) const {
        const float c00 = lerpf(ld(f0, b), ld(f1, b), al);
        const float c01 = lerpf(ld(f0, b + 1), ld(f1, b + 1), al);
        const float c10 = lerpf(ld(f0, b + xi), ld(f1, b + xi), al);
        const float c11 = lerpf(ld(f0, b + xi + 1), ld(f1, b + xi + 1), al);
        return lerpf(lerpf(c00, c01, fi), lerpf(c10, c11, fi), fj);
    }
    __device__ __forceinline__ float h(size_t c, float fi, float fj) const {
        return lerpf(lerpf(hf[c], hf[c + 1], fi), lerpf(hf[c + xi], hf[c + xi + 1], fi), fj);
    }
    __device__ __forceinline__ float zeta(size_t c, float fi, float fj, float al) const {
        return bil(zeta0, zeta1, c, fi, fj, al);
    }
    // velocity at rho level kI (u, v) and W levels kI, kI-1 (w), weight kF on kI-1
    __device__ __forceinline__ void velocity(int kI, float kF, size_t c, float fi, float fj, float al, float &uu,
                                             float &vv, float &ww) const {
        const size_t br = (size_t)(kI + s_rho - 1) * plane + c;
        uu = bil(u0, u1, br, fi, fj, al);
        vv = bil(v0, v1, br, fi, fj, al);
        const size_t bw = (size_t)(kI + s_w - 1) * plane + c;
        const float wt = bil(w0, w1, bw, fi, fj, al);
        const float wb = bil(w0, w1, bw - plane, fi, fj, al);
        ww = lerpf(wt, wb, kF);
    }
    __device__ __forceinline__ float akt(int kI, float kF, size_t c, float fi, float fj, float al) const {
        const size_t bw = (size_t)(kI + s_w - 1) * plane + c;
        return lerpf(bil(a0, a1, bw, fi, fj, al), bil(a0, a1, bw - plane, fi, fj, al), kF);
    }
};

__device__ __forceinline__ bool valid_cell(int j, int i, int eta, int xi) {
    return j >= 0 && i >= 0 && j + 1 < eta && i + 1 < xi;
}

// reflectDomain (NumericalHelpers.hpp:67-74), rare path, in FP64
__device__ double reflect_domain(double c, double m) {
    if (m <= 0) return 0;
    double r = fmod(c, 2.0 * m);
    if (r < 0) r += 2.0 * m;
    if (r >= m) r = 2.0 * m - r;
    if (r >= m) r = nextafter(m, 0.0);
    return r;
}

// Split a coordinate held as cell + offset after a displacement d (in cells):
// returns the new cell, offset in [0,1).
__device__ __forceinline__ void advance(int cell, float off, float d, int &ncell, float &noff) {
    const float x = off + d;
    const float fl = floorf(x);
    ncell = cell + (int)fl;
    noff = x - fl;
    if (noff >= 1.0f) {   // x just below an integer rounds up
        noff -= 1.0f;
        ncell += 1;
    }
}

template <class T>
__global__ __launch_bounds__(256) void move_fast_kernel(MoveParams P, FastSampler<T> S, SoA st, int n,
                                                        unsigned long long *stepCounter) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned steps = 0;
    if (idx < n) {
        const uint64_t id = st.id[idx];
        double kd = st.k[idx], jd = st.j[idx], id_ = st.i[idx];
        double health = st.health[idx], age = st.age[idx];
        const double ptime = st.time[idx];
        const float kLL = (float)P.kLowerLimit;
        const int eta = P.eta, xi = P.xi;

        // pre-step closures (Particle.cpp:134-176)
        if (kd > 0) {
            if (P.upper == WC_CLOSURE_CONSTRAINT) kd = 0;
            else if (P.upper == WC_CLOSURE_KILL) health = -1;
            else if (P.upper == WC_CLOSURE_REFLECTION) kd = -kd;
        }
        if (kd < P.kLowerLimit) {
            if (P.lower == WC_CLOSURE_CONSTRAINT) kd = P.kLowerLimit;
            else if (P.lower == WC_CLOSURE_KILL) health = -1;
            else if (P.lower == WC_CLOSURE_REFLECTION) kd = 2.0 * P.kLowerLimit - kd;
        }
        // register state: cell + FP32 offsets
        const double jfl = floor(jd), ifl = floor(id_);
        bool inside = jfl >= 0 && ifl >= 0 && jfl + 1 < eta && ifl + 1 < xi && kd == kd;
        int jc = inside ? (int)jfl : 0, ic = inside ? (int)ifl : 0;
        float fj = inside ? (float)(jd - jfl) : 0.f, fi = inside ? (float)(id_ - ifl) : 0.f;
        if (fj >= 1.0f) { fj -= 1.0f; ++jc; }
        if (fi >= 1.0f) { fi -= 1.0f; ++ic; }
        float k = (float)kd;
        bool healthFresh = true;   // `health` holds the current value (else derive from age)
        bool changed = false;

        const float dir = (float)P.direction;
        const float sv = (float)P.sv, shore = (float)P.shoreLimit;
        const uint32_t seedLo = (uint32_t)P.seed, seedHi = (uint32_t)(P.seed >> 32);
        const uint32_t ik = (uint32_t)P.intervalKey ^ (uint32_t)((uint64_t)P.intervalKey >> 32);
        const double L = P.intervalLength;
        double elapsed = P.elapsed0;

        while (elapsed < L) {
            const double stepDt = fmin(P.dti, L - elapsed);
            const double physicalTime = P.intervalStart + P.dirInt * elapsed;
            if (P.dirInt > 0 && ptime > physicalTime) {           // not yet emitted (Particle.cpp:211-215)
                elapsed = fmin(L, ptime - P.intervalStart);
                continue;
            }
            if (healthFresh ? (health < P.survprob) : (age > P.ageKill)) {   // :218-223
                health = -1;
                healthFresh = true;
                changed = true;
                break;
            }
            if (!inside || !(k == k)) {                           // open boundary (:241-248)
                health = -1;
                healthFresh = true;
                changed = true;
                break;
            }
            ++steps;
            changed = true;
            const float alpha = L == 0 ? 0.f : (float)((elapsed + 0.5 * stepDt) / L);
            const float alphaStart = L == 0 ? 0.f : (float)(elapsed / L);
            const float dt = (float)stepDt;
            const int kI = (int)ceilf(k);
            const float kF = (float)kI - k;
            const size_t c = (size_t)jc * xi + ic;

            const float hc = S.h(c, fi, fj) + S.zeta(c, fi, fj, alpha);
            bool kill = false;
            if (hc <= shore && P.horizontal == WC_CLOSURE_KILL) kill = true;

            float uu, vv, ww;
            S.velocity(kI, kF, c, fi, fj, P.midpoint ? alphaStart : alpha, uu, vv, ww);
            float rx = 0, ry = 0, rz = 0;
            if (P.randomOn) {
                const float aa = S.akt(kI, kF, c, fi, fj, alpha);
                const float sigmaDepth = (float)P.sigma * (1.0f - k / kLL);
                const uint32_t substep = (uint32_t)floor(elapsed / P.dti);
                const Philox4 r = philox4x32_10(substep, ik, (uint32_t)id, (uint32_t)(id >> 32), seedLo, seedHi);
                const float u1 = ((float)(r.v[0] >> 8) + 0.5f) * 0x1.0p-24f;
                const float u2 = (float)(r.v[1] >> 8) * 0x1.0p-24f;
                const float u3 = ((float)(r.v[2] >> 8) + 0.5f) * 0x1.0p-24f;
                const float u4 = (float)(r.v[3] >> 8) * 0x1.0p-24f;
                const float ra = sqrtf(-2.0f * __logf(u1)), rb = sqrtf(-2.0f * __logf(u3));
                float s1, c1, s2, c2;
                __sincosf(6.2831853071795865f * u2, &s1, &c1);
                __sincosf(6.2831853071795865f * u4, &s2, &c2);
                const float scale = sqrtf((float)(stepDt / P.dti)) * sigmaDepth;
                rx = scale * ra * c1;
                ry = scale * ra * s1;
                rz = scale * rb * c2 * aa * (float)P.crid;
            }
            const float idx_ = S.idxf[c], idy_ = S.idyf[c];
            const float zdist = hc * S.depthf[kI + P.s_w - 2];
            float dxl = dir * uu * dt, dyl = dir * vv * dt, dzl = dir * (sv + ww) * dt;

            if (P.midpoint) {                                      // Particle.cpp:585-641
                float kHalf = dir * (sv + ww) * 0.5f * dt;
                if (fabsf(kHalf) > zdist) kHalf = copysignf(fabsf(zdist), kHalf);
                int jm, im;
                float fjm, fim;
                advance(jc, fj, dir * vv * 0.5f * dt * idy_, jm, fjm);
                advance(ic, fi, dir * uu * 0.5f * dt * idx_, im, fim);
                const float kMid = fminf(0.0f, fmaxf(kLL, k + kHalf / zdist));
                bool midValid = valid_cell(jm, im, eta, xi);
                float hcMid = 0;
                size_t cm = 0;
                if (midValid) {
                    cm = (size_t)jm * xi + im;
                    hcMid = S.h(cm, fim, fjm) + S.zeta(cm, fim, fjm, alpha);
                    midValid = hcMid > shore && hcMid > 0;
                }
                if (midValid) {
                    const int kMI = (int)ceilf(kMid);
                    const float kMF = (float)kMI - kMid;
                    float um, vm, wm;
                    S.velocity(kMI, kMF, cm, fim, fjm, alpha, um, vm, wm);
                    // leap in metres of the start cell: u* dt dx/dx*
                    dxl = dir * um * dt * (S.idxf[cm] / idx_);
                    dyl = dir * vm * dt * (S.idyf[cm] / idy_);
                    dzl = dir * (sv + wm) * dt * (zdist / (hcMid * S.depthf[kMI + P.s_w - 2]));
                } else {
                    float ue, ve, we;
                    S.velocity(kI, kF, c, fi, fj, alpha, ue, ve, we);
                    dxl = dir * ue * dt;
                    dyl = dir * ve * dt;
                    dzl = dir * (sv + we) * dt;
                }
            }
            const float xleap = dxl + rx, yleap = dyl + ry;
            float zleap = dzl + rz;
            if (fabsf(zleap) > zdist) zleap = copysignf(fabsf(zdist), zleap);   // :644-648
            int jn, in_;
            float fjn, fin;
            advance(jc, fj, yleap * idy_, jn, fjn);
            advance(ic, fi, xleap * idx_, in_, fin);
            float kdet = k + zleap / zdist;
            if (kdet > 0.f) {                                      // :665-704
                if (P.upper == WC_CLOSURE_CONSTRAINT) kdet = 0.f;
                else if (P.upper == WC_CLOSURE_KILL) kill = true;
                else if (P.upper == WC_CLOSURE_REFLECTION) kdet = -kdet;
            }
            if (kdet < kLL) {
                if (P.lower == WC_CLOSURE_CONSTRAINT) kdet = kLL;
                else if (P.lower == WC_CLOSURE_KILL) kill = true;
                else if (P.lower == WC_CLOSURE_REFLECTION) kdet = 2.0f * kLL - kdet;
            }
            if (valid_cell(jn, in_, eta, xi)) {                    // :709-776
                const size_t cn = (size_t)jn * xi + in_;
                const float hcdet = S.h(cn, fin, fjn) + S.zeta(cn, fin, fjn, alpha);
                if (hcdet <= shore) {
                    if (P.horizontal == WC_CLOSURE_CONSTRAINT) {
                        jn = jc; fjn = fj; in_ = ic; fin = fi;
                    } else if (P.horizontal == WC_CLOSURE_KILL) {
                        kill = true;
                    } else if (P.horizontal == WC_CLOSURE_REFLECTION) {
                        // reflectCell (NumericalHelpers.hpp:62-65) relative to the old cell:
                        // behind -> old cell + |old - candidate|; ahead -> candidate cell - fraction
                        if (in_ < ic) advance(ic, 0.0f, fabsf(fi - ((float)(in_ - ic) + fin)), in_, fin);
                        else if (in_ > ic) advance(in_, 0.0f, -fin, in_, fin);
                        if (jn < jc) advance(jc, 0.0f, fabsf(fj - ((float)(jn - jc) + fjn)), jn, fjn);
                        else if (jn > jc) advance(jn, 0.0f, -fjn, jn, fjn);
                    }
                }
                jc = jn; fj = fjn; ic = in_; fi = fin; k = kdet;
            } else {                                               // :777-791
                if (P.horizontal == WC_CLOSURE_CONSTRAINT) {
                    k = kdet;
                } else if (P.horizontal == WC_CLOSURE_KILL) {
                    kill = true;
                } else if (P.horizontal == WC_CLOSURE_REFLECTION) {
                    const double ri = reflect_domain((double)in_ + (double)fin, xi - 1);
                    const double rj = reflect_domain((double)jn + (double)fjn, eta - 1);
                    const double rif = floor(ri), rjf = floor(rj);
                    ic = (int)rif; fi = (float)(ri - rif);
                    jc = (int)rjf; fj = (float)(rj - rjf);
                    k = kdet;
                }
            }
            inside = valid_cell(jc, ic, eta, xi) && fj == fj && fi == fi;
            if (kill) {
                health = -1;
                healthFresh = true;
            } else if (!(healthFresh && health <= 0)) {            // :797-804
                age = age + stepDt;
                healthFresh = false;
            }
            elapsed += stepDt;
        }
        if (!healthFresh) health = exp(-age / P.tau0);
        if (changed) {
            st.k[idx] = (double)k;
            st.j[idx] = (double)jc + (double)fj;
            st.i[idx] = (double)ic + (double)fi;
            st.health[idx] = health;
            st.age[idx] = age;
        } else if (kd != st.k[idx] || health != st.health[idx]) {  // pre-step closures only
            st.k[idx] = kd;
            st.health[idx] = health;
        }
    }
    if (stepCounter) wave_add_u64(stepCounter, steps);
}

template <class T>
hipError_t launch_t(const MoveParams &P, const void *const *pr, const GridPtrs &G, const SoA &s, int n,
                    unsigned long long *ctr, hipStream_t stream) {
    FastSampler<T> S;
    S.zeta0 = (const T *)pr[0]; S.zeta1 = (const T *)pr[1];
    S.u0 = (const T *)pr[2]; S.u1 = (const T *)pr[3];
    S.v0 = (const T *)pr[4]; S.v1 = (const T *)pr[5];
    S.w0 = (const T *)pr[6]; S.w1 = (const T *)pr[7];
    S.a0 = (const T *)pr[8]; S.a1 = (const T *)pr[9];
    S.hf = G.hf; S.idxf = G.idxf; S.idyf = G.idyf; S.depthf = G.depthf;
    S.xi = P.xi;
    S.plane = (size_t)P.eta * P.xi;
    S.s_rho = P.s_rho;
    S.s_w = P.s_w;
    const int block = 256;
    move_fast_kernel<T><<<(n + block - 1) / block, block, 0, stream>>>(P, S, s, n, ctr);
    return hipGetLastError();
}

__global__ void to_bf16_kernel(const float *__restrict__ src, uint16_t *__restrict__ dst, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const uint32_t b = __float_as_uint(src[q]);
        // round to nearest even (finite inputs)
        dst[q] = (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
    }
}

__global__ void to_f16_kernel(const float *__restrict__ src, _Float16 *__restrict__ dst, size_t n) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = (_Float16)src[q];
}

}  // namespace

hipError_t launch_move_fast_soa(int store, const MoveParams &P, const void *const *pair, const GridPtrs &G,
                                const SoA &s, int n, unsigned long long *ctr, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (store == WC_STORE_F16) return launch_t<_Float16>(P, pair, G, s, n, ctr, stream);
    if (store == WC_STORE_BF16) return launch_t<uint16_t>(P, pair, G, s, n, ctr, stream);
    return launch_t<float>(P, pair, G, s, n, ctr, stream);
}

hipError_t launch_convert_store(int store, const float *src, void *dst, size_t n, hipStream_t stream) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return hipSuccess;
    if (store == WC_STORE_BF16) to_bf16_kernel<<<(int)blocks, 256, 0, stream>>>(src, (uint16_t *)dst, n);
    else if (store == WC_STORE_F16) to_f16_kernel<<<(int)blocks, 256, 0, stream>>>(src, (_Float16 *)dst, n);
    return hipGetLastError();
}

}  // namespace wc
