/*
 * wc_oracle.c -- CPU restatement of WaComM++'s particle move + gridding.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): see wc_oracle.h for what may use it
 * and how each function is pinned.  Build: oracle/Makefile
 * (gcc -O2 -fopenmp -ffp-contract=off; never -ffast-math).
 *
 * WCO_PLANT=n builds a deliberately wrong variant (one plausible mistake per n)
 * that the CPU tests use to show their pins catch such mistakes.  The default
 * build has WCO_PLANT=0.
 */
#include "wc_oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>

#ifndef WCO_PLANT
#define WCO_PLANT 0
#endif

/* std::max / std::min exactly as <algorithm> defines them (the first argument
 * is returned when the comparison is false), so signed zeros match. */
static double std_max(double a, double b) { return (a < b) ? b : a; }
static double std_min(double a, double b) { return (b < a) ? b : a; }

/* ---- keyed RNG: NumericalHelpers.hpp:84-120 (ordered-splitmix64-v2) ------ */

uint64_t wco_mix(uint64_t value) {                      /* NumericalHelpers.hpp:84-89 */
    value += 0x9e3779b97f4a7c15ULL;
    value = (value ^ (value >> 30)) * 0xbf58476d1ce4e5b9ULL;
    value = (value ^ (value >> 27)) * 0x94d049bb133111ebULL;
    return value ^ (value >> 31);
}

static uint64_t combine(uint64_t state, uint64_t coordinate) {   /* NumericalHelpers.hpp:100-102 */
    return wco_mix(state ^ coordinate);
}

double wco_uniform(uint64_t seed, uint64_t particle, int64_t interval, uint64_t substep,
                   uint64_t component) {                 /* NumericalHelpers.hpp:104-112 */
    uint64_t key = wco_mix(seed);
    key = combine(key, particle);
    key = combine(key, (uint64_t)interval);
    key = combine(key, substep);
    key = combine(key, component);
    return (double)(key >> 11) * 0x1.0p-53;
}

double wco_normal(uint64_t seed, uint64_t particle, int64_t interval, uint64_t substep,
                  uint64_t component) {                  /* NumericalHelpers.hpp:114-120 */
    double u1 = std_max(wco_uniform(seed, particle, interval, substep, component * 2),
                        2.2250738585072014e-308);                       /* DBL_MIN */
    double u2 = wco_uniform(seed, particle, interval, substep, component * 2 + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.28318530717958647692 * u2);
}

/* ---- drift side helpers: DriftModel.hpp:110-160 --------------------------- */

double wco_jibe_step_probability(double hourlyProbability, double elapsedSeconds) {   /* :156-160 */
    if (hourlyProbability <= 0 || elapsedSeconds <= 0) return 0;
    if (hourlyProbability >= 1) return 1;
    return 1 - exp(log(1 - hourlyProbability) * elapsedSeconds / 3600);
}

int wco_sample_drift_side(uint64_t seed, uint64_t particle, double rightProbability) {   /* :146-154 */
    double u = wco_uniform(seed, particle, WCO_LEEWAY_SIDE_INTERVAL, 0, 0);
    return u < rightProbability ? 1 : -1;                 /* RIGHT : LEFT */
}

/* ---- numerical helpers: NumericalHelpers.hpp:45-74 ----------------------- */

double wco_restart_elapsed(double s, double e, double checkpoint) {   /* :45-52 */
    if (e > s) {
        if (e <= checkpoint) return NAN;
        return std_min(e - s, std_max(0.0, checkpoint - s));
    }
    if (e >= checkpoint) return NAN;
    return std_min(s - e, std_max(0.0, s - checkpoint));
}

double wco_reflect_cell(double oldCoordinate, int oldCell, double candidate, int candidateCell) {
    /* NumericalHelpers.hpp:62-65 */
    if (candidateCell < oldCell) candidate = oldCell + fabs(oldCoordinate - candidate);
    else if (candidateCell > oldCell) candidate = candidateCell - fmod(candidate, 1.0);
    return candidate;
}

double wco_reflect_domain(double coordinate, double maximum) {        /* :67-74 */
    if (maximum <= 0) return 0;
    double reflected = fmod(coordinate, 2.0 * maximum);
    if (reflected < 0) reflected += 2.0 * maximum;
    if (reflected >= maximum) reflected = 2.0 * maximum - reflected;
    if (reflected >= maximum) reflected = nextafter(maximum, 0.0);
    return reflected;
}

/* Fortran-style SIGN(a,b) = |a| * sgn(b) (Particle.cpp:833-841). */
static double sign_of(double a, double b) {
    double s = (double)((b > 0) - (b < 0));
    return fabs(a) * s;
}

/* validInterpolationCell (NumericalHelpers.hpp:12-14) evaluated on the floored
 * coordinate in double, so that NaN and out-of-int-range coordinates are
 * invalid (what the reference gets from (int) on x86-64) without C UB. */
static int valid_cell(double jf, double if_, int eta, int xi) {
    return jf >= 0.0 && if_ >= 0.0 && jf + 1.0 < (double)eta && if_ + 1.0 < (double)xi;
}

/* ---- the haversine cell metric: Particle.cpp:544-574 ---------------------- */

void wco_cell_metric(const wco_grid *g, int j, int i, double *dx, double *dy) {
    const int xi = g->xi;
    double dLat, dLon, dLatLon;
    /* Particle.cpp:548-559 */
    dLat = g->latRad[(size_t)(j + 1) * xi + i] - g->latRad[(size_t)j * xi + i];
    dLon = 0;
    dLatLon = pow(sin(0.5 * dLat), 2) +
              pow(sin(0.5 * dLon), 2) * cos(g->latRad[(size_t)(j + 1) * xi + i]) * cos(g->latRad[(size_t)j * xi + i]);
    *dy = 2.0 * atan2(pow(dLatLon, .5), pow(1.0 - dLatLon, .5)) * 6371000.0;
    /* Particle.cpp:563-574 */
    dLat = 0;
    dLon = g->lonRad[(size_t)j * xi + i + 1] - g->lonRad[(size_t)j * xi + i];
    dLatLon = pow(sin(0.5 * dLat), 2) +
              pow(sin(0.5 * dLon), 2) * cos(g->latRad[(size_t)j * xi + i + 1]) * cos(g->latRad[(size_t)j * xi + i]);
    *dx = 2.0 * atan2(pow(dLatLon, .5), pow(1.0 - dLatLon, .5)) * 6371000.0;
}

/* Haversine at the midpoint cell (Particle.cpp:616-621): same cells, with the
 * orthogonal term simply absent instead of multiplied by zero. */
static void cell_metric_mid(const wco_grid *g, int j, int i, double *dx, double *dy) {
    const int xi = g->xi;
    double mLat = g->latRad[(size_t)(j + 1) * xi + i] - g->latRad[(size_t)j * xi + i];
    double mLatLon = pow(sin(0.5 * mLat), 2);
    *dy = 2.0 * atan2(pow(mLatLon, .5), pow(1.0 - mLatLon, .5)) * 6371000.0;
    double mLon = g->lonRad[(size_t)j * xi + i + 1] - g->lonRad[(size_t)j * xi + i];
    mLatLon = pow(sin(0.5 * mLon), 2) * cos(g->latRad[(size_t)j * xi + i + 1]) * cos(g->latRad[(size_t)j * xi + i]);
    *dx = 2.0 * atan2(pow(mLatLon, .5), pow(1.0 - mLatLon, .5)) * 6371000.0;
}

/* ---- time-interpolated sampling: Particle.cpp:186-202 --------------------- */

typedef struct {
    const wco_grid *g;
    const wco_record *r0, *r1;
    size_t plane;        /* eta*xi */
} ctx_t;

/* (float)interpolateTime(f(t0), f(t1), a)  (NumericalHelpers.hpp:37-39) */
static float lerp_t(const float *f0, const float *f1, size_t idx, double a) {
#if WCO_PLANT == 1
    return (float)(a * (double)f0[idx] + (1.0 - a) * (double)f1[idx]);   /* planted: swapped weights */
#else
    return (float)((1.0 - a) * (double)f0[idx] + a * (double)f1[idx]);
#endif
}

static float zeta_at(const ctx_t *c, int j, int i, double a) {
    size_t idx = (size_t)j * c->g->xi + i;
    return lerp_t(c->r0->zeta, c->r1->zeta, idx, a);
}

/* level index of logical level k in a field of `levels` levels */
static size_t at3(const ctx_t *c, int k, int levels, int j, int i) {
    return (size_t)(k + levels - 1) * c->plane + (size_t)j * c->g->xi + i;
}

/* advectiveVelocity (Particle.cpp:352-475), passive particles only. */
static void advective_velocity(const ctx_t *c, int kI, double kF, int jI, double jF, int iI, double iF,
                               double a, float *uu, float *vv, float *ww) {
    const int sr = c->g->s_rho, sw = c->g->s_w;
    const float *u0 = c->r0->u, *u1 = c->r1->u, *v0 = c->r0->v, *v1 = c->r1->v;
    const float *w0 = c->r0->w, *w1 = c->r1->w;
#if WCO_PLANT == 2
    int ku = kI - 1;                                   /* planted: wrong rho level for u, v */
    if (ku < -sr + 1) ku = -sr + 1;
#else
    int ku = kI;
#endif
    float u1_ = lerp_t(u0, u1, at3(c, ku, sr, jI, iI), a) * (1.0 - iF) * (1.0 - jF);
    float u2_ = lerp_t(u0, u1, at3(c, ku, sr, jI + 1, iI), a) * (1.0 - iF) * jF;
    float u3_ = lerp_t(u0, u1, at3(c, ku, sr, jI + 1, iI + 1), a) * iF * jF;
    float u4_ = lerp_t(u0, u1, at3(c, ku, sr, jI, iI + 1), a) * iF * (1.0 - jF);
    *uu = u1_ + u2_ + u3_ + u4_;

    float v1_ = lerp_t(v0, v1, at3(c, ku, sr, jI, iI), a) * (1.0 - iF) * (1.0 - jF);
    float v2_ = lerp_t(v0, v1, at3(c, ku, sr, jI + 1, iI), a) * (1.0 - iF) * jF;
    float v3_ = lerp_t(v0, v1, at3(c, ku, sr, jI + 1, iI + 1), a) * iF * jF;
    float v4_ = lerp_t(v0, v1, at3(c, ku, sr, jI, iI + 1), a) * iF * (1.0 - jF);
    *vv = v1_ + v2_ + v3_ + v4_;

#if WCO_PLANT == 3
    const double kw = kF;                              /* planted: W-level weights swapped */
    const double kwl = 1.0 - kF;
#else
    const double kw = 1.0 - kF;
    const double kwl = kF;
#endif
    float w1_ = lerp_t(w0, w1, at3(c, kI, sw, jI, iI), a) * (1.0 - iF) * (1.0 - jF) * kw;
    float w2_ = lerp_t(w0, w1, at3(c, kI, sw, jI + 1, iI), a) * (1.0 - iF) * jF * kw;
    float w3_ = lerp_t(w0, w1, at3(c, kI, sw, jI + 1, iI + 1), a) * iF * jF * kw;
    float w4_ = lerp_t(w0, w1, at3(c, kI, sw, jI, iI + 1), a) * iF * (1.0 - jF) * kw;
    float w5_ = lerp_t(w0, w1, at3(c, kI - 1, sw, jI, iI), a) * (1.0 - iF) * (1.0 - jF) * kwl;
    float w6_ = lerp_t(w0, w1, at3(c, kI - 1, sw, jI + 1, iI), a) * (1.0 - iF) * jF * kwl;
    float w7_ = lerp_t(w0, w1, at3(c, kI - 1, sw, jI + 1, iI + 1), a) * iF * jF * kwl;
    float w8_ = lerp_t(w0, w1, at3(c, kI - 1, sw, jI, iI + 1), a) * iF * (1.0 - jF) * kwl;
    *ww = w1_ + w2_ + w3_ + w4_ + w5_ + w6_ + w7_ + w8_;
}

static double h_at(const wco_grid *g, int j, int i) { return g->h[(size_t)j * g->xi + i]; }

/* ---- Particle::move: Particle.cpp:71-824 ----------------------------------- */

void wco_move(const wco_config *cfg, const wco_grid *g, double intervalStart, double intervalEnd,
              const wco_record *rec0, const wco_record *rec1, wco_particle *pp) {
    wco_particle P = *pp;                                   /* Particle.cpp:77-78 */
    const int random = cfg->random;                         /* :81-105 */
    const double dti = cfg->dti, tau0 = cfg->tau0, survprob = cfg->survprob;
    const double crid = cfg->crid, sv = cfg->sv, sigma = cfg->sigma, shoreLimit = cfg->shoreLimit;
    const double health0 = 1.0;                             /* Particle.hpp:92 */

    double intervalLength = fabs(intervalEnd - intervalStart);   /* :115 */
    double elapsed = 0;
    if (!isnan(cfg->restartCheckpoint)) {                   /* :117-120 */
        elapsed = wco_restart_elapsed(intervalStart, intervalEnd, cfg->restartCheckpoint);
        if (isnan(elapsed)) return;
    }
    const int s_w = g->s_w;
    const double kLowerLimit = -s_w + 2;                    /* :125 */
    const int eta = g->eta, xi = g->xi;

    ctx_t c = { g, rec0, rec1, (size_t)eta * xi };

    if (P.k > 0) {                                          /* :134-153 */
        switch (cfg->upperClosure) {
            case WCO_CONSTRAINT: P.k = 0; break;
            case WCO_KILL: P.health = -1; break;
            case WCO_REFLECTION: P.k = -P.k; break;
        }
    }
    if (P.k < kLowerLimit) {                                /* :156-176 */
        switch (cfg->lowerClosure) {
            case WCO_CONSTRAINT: P.k = kLowerLimit; break;
            case WCO_KILL: P.health = -1; break;
            case WCO_REFLECTION: P.k = 2.0 * kLowerLimit - P.k; break;
        }
    }

    const double direction = cfg->trackingDirection;        /* :507 */
    const int midpoint = cfg->integrator == WCO_MIDPOINT;   /* :479 */
    /* random-key coordinates (Particle.cpp:523) */
    const int64_t intervalKey = llround(intervalStart < intervalEnd ? intervalStart : intervalEnd);

    while (elapsed < intervalLength) {                      /* :178 */
        double stepDt = std_min(dti, intervalLength - elapsed);   /* :180 */
        double physicalTime = intervalStart + cfg->trackingDirection * elapsed;   /* :181 */
        double alpha = intervalLength == 0 ? 0 : (elapsed + 0.5 * stepDt) / intervalLength;  /* :182 */
        double alphaStart = intervalLength == 0 ? 0 : elapsed / intervalLength;           /* :190 */

        if (cfg->trackingDirection > 0 && P.time > physicalTime) {   /* :211-215 */
            elapsed = std_min(intervalLength, P.time - intervalStart);
            continue;
        }
        if (P.health < survprob) { P.health = -1; break; }  /* :218-223 */

        double kd = ceil(P.k);                              /* :231-238 */
        double kF = ceil(P.k) - P.k;
        double jd = floor(P.j), id = floor(P.i);
        if (!valid_cell(jd, id, eta, xi)) { P.health = -1; break; }   /* :241-248 */
        /* Memory-safety guard: a level outside [kLowerLimit, 0] (only reachable
         * from non-finite k or a double closure overshoot) would make the
         * reference read outside u/v/w/akt; kill as cuda/kernel.cu:195 does. */
        if (!(kd <= 0.0 && kd >= kLowerLimit)) { P.health = -1; break; }
        const int kI = (int)kd, jI = (int)jd, iI = (int)id;
        double jF = P.j - jI, iF = P.i - iI;

        /* zeta and h at the particle (Particle.cpp:271-290) */
        float z1 = zeta_at(&c, jI, iI, alpha) * (1.0 - iF) * (1.0 - jF);
        float z2 = zeta_at(&c, jI + 1, iI, alpha) * (1.0 - iF) * jF;
        float z3 = zeta_at(&c, jI + 1, iI + 1, alpha) * iF * jF;
        float z4 = zeta_at(&c, jI, iI + 1, alpha) * iF * (1.0 - jF);
        float zz = z1 + z2 + z3 + z4;
        double h1 = h_at(g, jI, iI) * (1.0 - iF) * (1.0 - jF);
        double h2 = h_at(g, jI + 1, iI) * (1.0 - iF) * jF;
        double h3 = h_at(g, jI + 1, iI + 1) * iF * jF;
        double h4 = h_at(g, jI, iI + 1) * iF * (1.0 - jF);
        double hh = h1 + h2 + h3 + h4;
        double hc = hh + zz;

        if (hc <= shoreLimit) {                             /* :311-338 */
            if (cfg->horizontalClosure == WCO_KILL) P.health = -1;
            /* constraint and reflection: no action here */
        }

        /* first stage (Particle.cpp:479-481) */
        float uu, vv, ww;
#if WCO_PLANT == 4
        advective_velocity(&c, kI, kF, jI, jF, iI, iF, midpoint ? alphaStart : alphaStart, &uu, &vv, &ww);  /* planted: Euler samples at the substep start */
#else
        advective_velocity(&c, kI, kF, jI, jF, iI, iF, midpoint ? alphaStart : alpha, &uu, &vv, &ww);
#endif

        /* AKt at the particle (Particle.cpp:490-500) */
        const int sw = s_w;
        const float *a0 = rec0->akt, *a1 = rec1->akt;
        float aa1 = lerp_t(a0, a1, at3(&c, kI, sw, jI, iI), alpha) * (1.0 - iF) * (1.0 - jF) * (1.0 - kF);
        float aa2 = lerp_t(a0, a1, at3(&c, kI, sw, jI + 1, iI), alpha) * (1.0 - iF) * jF * (1.0 - kF);
        float aa3 = lerp_t(a0, a1, at3(&c, kI, sw, jI + 1, iI + 1), alpha) * iF * jF * (1.0 - kF);
        float aa4 = lerp_t(a0, a1, at3(&c, kI, sw, jI, iI + 1), alpha) * iF * (1.0 - jF) * (1.0 - kF);
        float aa5 = lerp_t(a0, a1, at3(&c, kI - 1, sw, jI, iI), alpha) * (1.0 - iF) * (1.0 - jF) * kF;
        float aa6 = lerp_t(a0, a1, at3(&c, kI - 1, sw, jI + 1, iI), alpha) * (1.0 - iF) * jF * kF;
        float aa7 = lerp_t(a0, a1, at3(&c, kI - 1, sw, jI + 1, iI + 1), alpha) * iF * jF * kF;
        float aa8 = lerp_t(a0, a1, at3(&c, kI - 1, sw, jI, iI + 1), alpha) * iF * (1.0 - jF) * kF;
        float aa = aa1 + aa2 + aa3 + aa4 + aa5 + aa6 + aa7 + aa8;

        /* deterministic leap (Particle.cpp:507-510) */
        double dxleap = direction * uu * stepDt;
        double dyleap = direction * vv * stepDt;
        double dzleap = direction * (sv + ww) * stepDt;

        /* random leap (Particle.cpp:513-533) */
        double rxleap = 0, ryleap = 0, rzleap = 0;
        if (random && (direction > 0 || cfg->backwardDiffusion)) {
#if WCO_PLANT == 5
            double sigmaDepth = sigma * (1 + P.k / kLowerLimit);        /* planted: wrong depth factor */
#else
            double sigmaDepth = sigma * (1 - P.k / kLowerLimit);
#endif
            uint64_t substep = (uint64_t)floor(elapsed / dti);
            double stochasticScale = sqrt(stepDt / dti);
            rxleap = sigmaDepth * wco_normal(cfg->randomSeed, P.id, intervalKey, substep, 0) * stochasticScale;
            ryleap = sigmaDepth * wco_normal(cfg->randomSeed, P.id, intervalKey, substep, 1) * stochasticScale;
            rzleap = sigmaDepth * wco_normal(cfg->randomSeed, P.id, intervalKey, substep, 2) * aa * crid * stochasticScale;
        }
        double xleap = dxleap + rxleap;                     /* :536-538 */
        double yleap = dyleap + ryleap;
        double zleap = dzleap + rzleap;

        double xdist, ydist;                                /* :544-579 */
        wco_cell_metric(g, jI, iI, &xdist, &ydist);
        double zdist = hc * g->depthIntervals[kI + s_w - 2];

        if (midpoint) {                                     /* :585-641 */
            double kHalfLeap = direction * (sv + ww) * 0.5 * stepDt;
            if (fabs(kHalfLeap) > zdist) kHalfLeap = sign_of(zdist, kHalfLeap);
            double iMid = P.i + direction * uu * 0.5 * stepDt / xdist;
            double jMid = P.j + direction * vv * 0.5 * stepDt / ydist;
            double kMid = std_min(0.0, std_max(kLowerLimit, P.k + kHalfLeap / zdist));
            double jMd = floor(jMid), iMd = floor(iMid);
            int midValid = valid_cell(jMd, iMd, eta, xi);
            double hcMid = 0;
            int iMI = 0, jMI = 0;
            if (midValid) {
                iMI = (int)iMd; jMI = (int)jMd;
                double iMF = iMid - iMI, jMF = jMid - jMI;
                hcMid = h_at(g, jMI, iMI) * (1.0 - iMF) * (1.0 - jMF) + h_at(g, jMI + 1, iMI) * (1.0 - iMF) * jMF +
                        h_at(g, jMI + 1, iMI + 1) * iMF * jMF + h_at(g, jMI, iMI + 1) * iMF * (1.0 - jMF) +
                        zeta_at(&c, jMI, iMI, alpha) * (1.0 - iMF) * (1.0 - jMF) +
                        zeta_at(&c, jMI + 1, iMI, alpha) * (1.0 - iMF) * jMF +
                        zeta_at(&c, jMI + 1, iMI + 1, alpha) * iMF * jMF +
                        zeta_at(&c, jMI, iMI + 1, alpha) * iMF * (1.0 - jMF);
                midValid = hcMid > shoreLimit && hcMid > 0;
            }
            if (midValid) {
                double iMF = iMid - iMI, jMF = jMid - jMI;
                double kMd = ceil(kMid);
                int kMI = (int)kMd;
                double kMF = ceil(kMid) - kMid;
                float um, vm, wm;
                advective_velocity(&c, kMI, kMF, jMI, jMF, iMI, iMF, alpha, &um, &vm, &wm);
                double xdistMid, ydistMid;
                cell_metric_mid(g, jMI, iMI, &xdistMid, &ydistMid);
                double zdistMid = hcMid * g->depthIntervals[kMI + s_w - 2];
#if WCO_PLANT == 6
                dxleap = direction * um * stepDt;           /* planted: midpoint metric not rescaled */
                dyleap = direction * vm * stepDt;
                dzleap = direction * (sv + wm) * stepDt;
                (void)xdistMid; (void)ydistMid; (void)zdistMid;
#else
                dxleap = direction * um * stepDt * xdist / xdistMid;
                dyleap = direction * vm * stepDt * ydist / ydistMid;
                dzleap = direction * (sv + wm) * stepDt * zdist / zdistMid;
#endif
            } else {
                float ue, ve, we;
                advective_velocity(&c, kI, kF, jI, jF, iI, iF, alpha, &ue, &ve, &we);
                dxleap = direction * ue * stepDt;
                dyleap = direction * ve * stepDt;
                dzleap = direction * (sv + we) * stepDt;
            }
            xleap = dxleap + rxleap;
            yleap = dyleap + ryleap;
            zleap = dzleap + rzleap;
        }

        if (fabs(zleap) > zdist) zleap = sign_of(zdist, zleap);   /* :644-648 */

        double jdet = P.j + yleap / ydist;                  /* :651-657 */
        double idet = P.i + xleap / xdist;
        double kdet = P.k + zleap / zdist;

        if (kdet > 0.) {                                    /* :665-683 */
            switch (cfg->upperClosure) {
                case WCO_CONSTRAINT: kdet = 0; break;
                case WCO_KILL: P.health = -1; break;
                case WCO_REFLECTION: kdet = -kdet; break;
            }
        }
        if (kdet < kLowerLimit) {                           /* :686-704 */
            switch (cfg->lowerClosure) {
                case WCO_CONSTRAINT: kdet = kLowerLimit; break;
                case WCO_KILL: P.health = -1; break;
                case WCO_REFLECTION: kdet = 2.0 * kLowerLimit - kdet; break;
            }
        }

        double jdd = floor(jdet), idd = floor(idet);        /* :709-713 */
        if (valid_cell(jdd, idd, eta, xi)) {
            int jdetI = (int)jdd, idetI = (int)idd;
            double idetF = idet - idetI;                    /* :716-739 */
            double jdetF = jdet - jdetI;
            z1 = zeta_at(&c, jdetI, idetI, alpha) * (1.0 - idetF) * (1.0 - jdetF);
            z2 = zeta_at(&c, jdetI + 1, idetI, alpha) * (1.0 - idetF) * jdetF;
            z3 = zeta_at(&c, jdetI + 1, idetI + 1, alpha) * idetF * jdetF;
            z4 = zeta_at(&c, jdetI, idetI + 1, alpha) * idetF * (1.0 - jdetF);
            float zzdet = z1 + z2 + z3 + z4;
            h1 = h_at(g, jdetI, idetI) * (1.0 - idetF) * (1.0 - jdetF);
            h2 = h_at(g, jdetI + 1, idetI) * (1.0 - idetF) * jdetF;
            h3 = h_at(g, jdetI + 1, idetI + 1) * idetF * jdetF;
            h4 = h_at(g, jdetI, idetI + 1) * idetF * (1.0 - jdetF);
            double hhdet = h1 + h2 + h3 + h4;
            double hcdet = hhdet + zzdet;
            if (hcdet <= shoreLimit) {                      /* :753-771 */
                switch (cfg->horizontalClosure) {
                    case WCO_CONSTRAINT: idet = P.i; jdet = P.j; break;
                    case WCO_KILL: P.health = -1; break;
                    case WCO_REFLECTION:
                        idet = wco_reflect_cell(P.i, iI, idet, idetI);
                        jdet = wco_reflect_cell(P.j, jI, jdet, jdetI);
                        break;
                }
            }
            P.i = idet; P.j = jdet; P.k = kdet;             /* :774-776 */
        } else {                                            /* :777-791 */
            switch (cfg->horizontalClosure) {
                case WCO_CONSTRAINT: P.k = kdet; break;
                case WCO_KILL: P.health = -1; break;
                case WCO_REFLECTION:
                    P.i = wco_reflect_domain(idet, xi - 1);
                    P.j = wco_reflect_domain(jdet, eta - 1);
                    P.k = kdet;
                    break;
            }
        }

        if (P.health > 0) {                                 /* :797-804 */
            P.age = P.age + stepDt;
            P.health = health0 * exp(-P.age / tau0);
            if (cfg->leewayJibeProbabilityHourly > 0 && P.driftObjectType != 0) {   /* :805-813 */
                uint64_t substep = (uint64_t)floor(elapsed / dti);
                double probability = wco_jibe_step_probability(cfg->leewayJibeProbabilityHourly, stepDt);
                if (wco_uniform(cfg->randomSeed, P.id, intervalKey, substep, WCO_JIBE_COMPONENT) < probability)
                    P.driftSide = (int8_t)-P.driftSide;
            }
        }
        elapsed += stepDt;                                  /* :821 */
    }
    *pp = P;                                                /* :823 */
}

void wco_move_all(const wco_config *cfg, const wco_grid *g, double intervalStart, double intervalEnd,
                  const wco_record *rec0, const wco_record *rec1, wco_particle *ps, long n, int nthreads) {
    /* Wacomm.cpp:501-525: omp parallel for schedule(dynamic, 64) */
    if (nthreads > 0) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
        for (long idx = 0; idx < n; idx++) wco_move(cfg, g, intervalStart, intervalEnd, rec0, rec1, &ps[idx]);
    } else {
#pragma omp parallel for schedule(dynamic, 64)
        for (long idx = 0; idx < n; idx++) wco_move(cfg, g, intervalStart, intervalEnd, rec0, rec1, &ps[idx]);
    }
}

void wco_move_window(const wco_config *cfg, const wco_grid *g, int ocean_time_idx, int T,
                     const double *oceanTime, const float *zeta, const float *u, const float *v,
                     const float *w, const float *akt, wco_particle *ps, long n, int nthreads) {
    /* Particle.cpp:110-115 and :183-184 */
    int dir = cfg->trackingDirection;
    int nextIdx = ocean_time_idx + dir;
    double intervalStart = oceanTime[ocean_time_idx];
    double intervalEnd = (nextIdx >= 0 && nextIdx < T) ? oceanTime[nextIdx] : intervalStart + dir * cfg->deltat;
    if (nextIdx < 0 || nextIdx >= T) nextIdx = ocean_time_idx;
    size_t plane = (size_t)g->eta * g->xi;
    wco_record r0 = { zeta + plane * ocean_time_idx, u + plane * g->s_rho * ocean_time_idx,
                      v + plane * g->s_rho * ocean_time_idx, w + plane * g->s_w * ocean_time_idx,
                      akt + plane * g->s_w * ocean_time_idx };
    wco_record r1 = { zeta + plane * nextIdx, u + plane * g->s_rho * nextIdx, v + plane * g->s_rho * nextIdx,
                      w + plane * g->s_w * nextIdx, akt + plane * g->s_w * nextIdx };
    wco_move_all(cfg, g, intervalStart, intervalEnd, &r0, &r1, ps, n, nthreads);
}

/* ---- Concentration::evaluate: Concentration.hpp:16-44 -------------------- */

void wco_grid_count(const wco_particle *ps, long n, int s_rho, int eta, int xi, uint32_t *counts) {
    for (long idx = 0; idx < n; idx++) {
        const wco_particle *p = &ps[idx];
        if (!(p->health > 0)) continue;                     /* Particle.cpp:65-67 isAlive */
#if WCO_PLANT == 7
        double kr = floor(p->k), jr = floor(p->j), ir = floor(p->i);   /* planted: floor, not round */
#else
        double kr = round(p->k), jr = round(p->j), ir = round(p->i);   /* :31-33, half away from zero */
#endif
        /* (int) of a value outside int range is UB in C; such values fail the
         * bounds test in the reference too (x86-64 yields INT_MIN). */
        if (!(jr >= 0 && jr < eta && ir >= 0 && ir < xi && kr >= -s_rho + 1 && kr <= 0)) continue;   /* :36 */
        int k = (int)kr, j = (int)jr, i = (int)ir;
        counts[((size_t)(k + s_rho - 1) * eta + j) * xi + i] += 1;   /* :39-40 */
    }
}

/* ---- sourceEmissionTimes: SourceEmission.hpp:38-79 ------------------------ */

long wco_emission_times(int mode, double value, double sourceStart, double sourceEnd,
                        double intervalStart, double intervalEnd, double *times, long cap) {
    long n = 0;
    if (!isfinite(value) || value < 0 || !isfinite(sourceStart) || !isfinite(intervalStart) ||
        !isfinite(intervalEnd) || intervalEnd <= intervalStart)
        return -1;                                          /* :42-44 */
    if (value == 0) return 0;                               /* :46 */
    if (mode == 2) {                                        /* SINGLE_PULSE :48-54 */
        if (floor(value) != value) return -1;
        if (sourceStart >= intervalStart && sourceStart < intervalEnd) {
            long cnt = (long)value;
            if (cnt > cap) return -2;
            for (long q = 0; q < cnt; q++) times[q] = sourceStart;
            return cnt;
        }
        return 0;
    }
    if (mode == 1) {                                        /* FORCING_INTERVAL_BATCH :56-66 */
        if (floor(value) != value) return -1;
        long long record = llround(intervalStart);
        if (record >= llround(sourceStart) && (sourceEnd < 0 || record <= llround(sourceEnd))) {
            long cnt = (long)value;
            if (cnt > cap) return -2;
            for (long q = 0; q < cnt; q++) times[q] = intervalStart;
            return cnt;
        }
        return 0;
    }
    /* UNIFORM_RATE :68-78 */
    double activeStart = std_max(intervalStart, sourceStart);
    double activeEnd = sourceEnd < 0 ? intervalEnd : std_min(intervalEnd, sourceEnd);
    if (activeEnd <= activeStart) return 0;
    double spacing = 3600.0 / value;
    long long first = (long long)ceil((activeStart - sourceStart) / spacing - 0.5);
    if (first < 0) first = 0;
    for (long long index = first;; index++) {
        double t = sourceStart + (index + 0.5) * spacing;
        if (t >= activeEnd) break;
        if (t >= activeStart) {
            if (n >= cap) return -2;
            times[n++] = t;
        }
    }
    return n;
}

/* ---- ROMS normalisation: ROMSAdapter.cpp:29-49, 56-300, 309-368, 448-466 */

#define WCO_NC_FILL_FLOAT 9.9692099683868690e+36f    /* netcdf.h NC_FILL_FLOAT */
#define WCO_NC_FILL_DOUBLE 9.9692099683868690e+36    /* netcdf.h NC_FILL_DOUBLE */

/* validRomsValue (ROMSAdapter.cpp:45-49) over the attribute list plus the
 * two NetCDF fill values romsMissingValues appends (:40-41). */
static int roms_valid(double value, const double *missing, int n) {
    if (!isfinite(value)) return 0;
    for (int q = 0; q < n; q++)
        if (value == missing[q]) return 0;
    if (value == (double)WCO_NC_FILL_FLOAT || value == WCO_NC_FILL_DOUBLE) return 0;
    return 1;
}

int wco_roms_normalize(const wco_roms_grid *g, const float *zeta, const float *u, const float *v, const float *w,
                       const float *akt, float *ZETA, float *U, float *V, float *W, float *AKT) {
    const int s_rho = g->s_rho, s_w = g->s_rho + 1, eta = g->eta, xi = g->xi;
    const int eta_u = g->eta_u, xi_u = g->xi_u, eta_v = g->eta_v, xi_v = g->xi_v;
    const size_t plane = (size_t)eta * xi;
    const double *mr = g->mask_rho, *mu = g->mask_u, *mv = g->mask_v;
    /* input checks of :146-166 */
    for (size_t c = 0; c < plane; c++) {
        if (mr[c] != 0.0 && mr[c] != 1.0) return -1;
        if (!g->earth_relative && mr[c] > 0.0 && !roms_valid(g->angle[c], g->missing_angle, g->n_missing_angle))
            return -2;
    }
    for (int j = 0; j < eta_u; j++)
        for (int i = 0; i < xi_u; i++) {
            const double m = mu[(size_t)j * xi_u + i];
            if (m != 0.0 && m != 1.0) return -1;
            if (m > 0.0)
                for (int k = 0; k < s_rho; k++)
                    if (!roms_valid(u[((size_t)k * eta_u + j) * xi_u + i], g->missing_u, g->n_missing_u)) return -3;
        }
    for (int j = 0; j < eta_v; j++)
        for (int i = 0; i < xi_v; i++) {
            const double m = mv[(size_t)j * xi_v + i];
            if (m != 0.0 && m != 1.0) return -1;
            if (m > 0.0)
                for (int k = 0; k < s_rho; k++)
                    if (!roms_valid(v[((size_t)k * eta_v + j) * xi_v + i], g->missing_v, g->n_missing_v)) return -4;
        }
    /* uv2rho :309-368: float sums over wet faces, float / int average */
    for (int k = 0; k < s_rho; k++)
        for (int j = 0; j < eta; j++)
            for (int i = 0; i < xi; i++) {
                const size_t o = (size_t)k * plane + (size_t)j * xi + i;
                float usum = 0.0f, vsum = 0.0f;
                int ucount = 0, vcount = 0;
                if (mr[(size_t)j * xi + i] > 0.0) {
                    if (j < eta_u && i < xi_u && mu[(size_t)j * xi_u + i] > 0.0) {
                        usum += u[((size_t)k * eta_u + j) * xi_u + i]; ucount++;
                    }
                    if (j < eta_u && i > 0 && i - 1 < xi_u && mu[(size_t)j * xi_u + i - 1] > 0.0) {
                        usum += u[((size_t)k * eta_u + j) * xi_u + i - 1]; ucount++;
                    }
                    if (j < eta_v && i < xi_v && mv[(size_t)j * xi_v + i] > 0.0) {
                        vsum += v[((size_t)k * eta_v + j) * xi_v + i]; vcount++;
                    }
                    if (j > 0 && j - 1 < eta_v && i < xi_v && mv[(size_t)(j - 1) * xi_v + i] > 0.0) {
                        vsum += v[((size_t)k * eta_v + j - 1) * xi_v + i]; vcount++;
                    }
                    U[o] = ucount > 0 ? usum / ucount : 0.0f;
                    V[o] = vcount > 0 ? vsum / vcount : 0.0f;
                } else {
                    U[o] = 0.0f;
                    V[o] = 0.0f;
                }
            }
    /* rotation :276-300: cosine/sine are 1/0 on land, east/north in double */
    int invalid = 0;
    for (int k = 0; k < s_rho; k++)
        for (size_t c = 0; c < plane; c++) {
            if (mr[c] <= 0.0) continue;
            const double cs = g->earth_relative ? 1.0 : cos(g->angle[c]);
            const double sn = g->earth_relative ? 0.0 : sin(g->angle[c]);
            const size_t o = (size_t)k * plane + c;
            const double gu = U[o], gv = V[o];
            const double east = g->earth_relative ? gu : gu * cs - gv * sn;
            const double north = g->earth_relative ? gv : gu * sn + gv * cs;
            if (!isfinite(east) || !isfinite(north) || fabs(east) > (double)FLT_MAX || fabs(north) > (double)FLT_MAX)
                invalid = 1;
            else {
                U[o] = (float)east;
                V[o] = (float)north;
            }
        }
    /* wakt2wakt :448-466 */
    for (int k = 0; k < s_w; k++)
        for (size_t c = 0; c < plane; c++) {
            const size_t o = (size_t)k * plane + c;
            W[o] = mr[c] > 0.0 ? w[o] : 0.0f;
            AKT[o] = mr[c] > 0.0 ? akt[o] : 0.0f;
        }
    /* zeta :254-264 */
    for (size_t c = 0; c < plane; c++)
        ZETA[c] = (mr[c] == 0 || zeta[c] >= 1e+36f) ? 0.0f : zeta[c];
    return invalid ? -5 : 0;
}

void wco_roms_static(int s_rho, int eta, int xi, const double *s_w_values, const double *lon, const double *lat,
                     const double *h, const double *mask_rho, double *lonRad, double *latRad, double *h_out,
                     double *depthIntervals) {
    const int s_w = s_rho + 1;
    /* DepthIntervals(k) = sW(k) - sW(k-1), k = -s_w+2..0 (:232-235); sW
       storage index k+s_w-1 */
    for (int k = -s_w + 2; k <= 0; k++)
        depthIntervals[k + s_w - 2] = s_w_values[k + s_w - 1] - s_w_values[k + s_w - 2];
    for (size_t c = 0; c < (size_t)eta * xi; c++) {
        h_out[c] = mask_rho[c] == 0 ? 0.0 : h[c];          /* :244-252 */
        lonRad[c] = 0.0174533 * lon[c];                     /* :268-275 */
        latRad[c] = 0.0174533 * lat[c];
    }
}

/* (int)floor(x) as x86-64 converts it: values outside int range and NaN give
 * INT_MIN (the reference's static_cast<int>(std::floor(x)),
 * NumericalHelpers.hpp:20-26). */
static int cell_of(double x) {
    const double f = floor(x);
    return (f >= -2147483648.0 && f < 2147483648.0) ? (int)f : INT32_MIN;
}

static int level_of(double x) {
    const double c = ceil(x);
    return (c >= -2147483648.0 && c < 2147483648.0) ? (int)c : INT32_MIN;
}

void wco_kji2deplatlon(double k, double j, double i, int s_w, int eta, int xi, const double *sW, const double *lat,
                       const double *lon, const double *h, double *dep, double *olat, double *olon) {
    const int kI = level_of(k);                          /* upperVerticalLevel */
    const double kF = ceil(k) - k;                       /* lowerVerticalWeight */
    const int jI = cell_of(j);
    const double jF = j - jI;
    const int iI = cell_of(i);
    const double iF = i - iI;
    /* validInterpolationCell (NumericalHelpers.hpp:12-14) and the level range
       (OceanModelAdapter.cpp:284-291) */
    if (!(jI >= 0 && iI >= 0 && (long)jI + 1 < eta && (long)iI + 1 < xi) || kI > 0 || kI <= -s_w + 1) {
        *dep = *olat = *olon = 1e37;
        return;
    }
    const size_t c00 = (size_t)jI * xi + iI, c10 = c00 + xi, c11 = c10 + 1, c01 = c00 + 1;
    /* :293-320, each term left to right, summed in order */
    *olon = lon[c00] * (1.0 - iF) * (1.0 - jF) + lon[c10] * (1.0 - iF) * jF + lon[c11] * iF * jF +
            lon[c01] * iF * (1.0 - jF);
    *olat = lat[c00] * (1.0 - iF) * (1.0 - jF) + lat[c10] * (1.0 - iF) * jF + lat[c11] * iF * jF +
            lat[c01] * iF * (1.0 - jF);
    const double hh = h[c00] * (1.0 - iF) * (1.0 - jF) + h[c10] * (1.0 - iF) * jF + h[c11] * iF * jF +
                      h[c01] * iF * (1.0 - jF);
    const double sLo = sW[kI - 1 + s_w - 1], sUp = sW[kI + s_w - 1];
    const double aDep = fabs(fabs(hh * sLo) - fabs(hh * sUp));           /* :315-318 */
    *dep = -(hh * fabs(sUp) + fabs(kF * aDep));                          /* :319 */
}

long wco_particle_records(const wco_particle *ps, long n, int s_w, int eta, int xi, const double *sW,
                          const double *lat, const double *lon, const double *h, long *src, double *dep,
                          double *olat, double *olon) {
    long r = 0;
    for (long q = 0; q < n; q++) {
        const wco_particle *p = &ps[q];
        if (p->j < 0 || p->i < 0 || p->k > 0) continue;                /* Particles.cpp:121-123 */
        double d, la, lo;
        wco_kji2deplatlon(p->k, p->j, p->i, s_w, eta, xi, sW, lat, lon, h, &d, &la, &lo);
        if (d == 1e37 || la == 1e37 || lo == 1e37) continue;          /* :128-130 */
        src[r] = q;
        dep[r] = d;
        olat[r] = la;
        olon[r] = lo;
        r++;
    }
    return r;
}
