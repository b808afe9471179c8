// Device-side data structures shared by the move and gridding kernels of
// libwacomm_hip.so (gfx950).  The parity and gridding sources are compiled with
// contraction off (-ffp-contract=off plus a file pragma) so that every
// rounding point of the reference's Particle::move is kept; the mixed-precision
// source fuses multiply-adds within each source expression (contract(on)).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wacomm_hip.h"

namespace wc {

// Scalars of one forcing interval, resolved on the host from config_data and
// the interval bounds (Particle.cpp:81-125, 507, 523).  Passed by value.
struct MoveParams {
    double dti, tau0, survprob, crid, sv, sigma, shoreLimit;
    double intervalStart;   // physical time of the interval start
    double intervalLength;  // |end - start|
    double elapsed0;        // restart-clipped start offset (Particle.cpp:116-120)
    double kLowerLimit;     // -s_w + 2
    double direction;       // trackingDirection as double
    double ageKill;         // exp(-age/tau0) < survprob  <=>  age > ageKill (mixed-precision path)
    double ageLazy;         // FP64 path: below it exp(-age/tau0) is surely >= survprob and > 0
    double invDti;          // 1 / dti (an estimate only: checked exactly before use)
    uint64_t seed;
    int64_t intervalKey;    // llround(min(start, end))
    double jibeHourly;      // leewayJibeProbabilityHourly
    int jibe;               // jibeHourly > 0: non-passive particles may flip their drift side
    int upper, lower, horizontal;
    int dirInt;
    int randomOn;           // random && (direction > 0 || backwardDiffusion)
    int midpoint;
    int s_w, s_rho, eta, xi;
};

// The two forcing records of the interval (record at the interval start and
// at its end), laid out per record as zeta [eta][xi], u/v [s_rho][eta][xi],
// w/akt [s_w][eta][xi].
template <class T>
struct RecordPair {
    const T *zeta0, *zeta1, *u0, *u1, *v0, *v1, *w0, *w1, *akt0, *akt1;
};

// Static grid: h [eta][xi] (FP64), depthIntervals [s_w-1], and either the
// per-cell haversine sizes dx, dy [eta][xi] or lon/lat to compute them.
// The mixed-precision path reads FP32 copies: h, 1/dx, 1/dy, depthIntervals.
struct GridPtrs {
    const double *h, *depth, *dx, *dy, *lonRad, *latRad;
    const float *hf, *idxf, *idyf, *depthf;
};

// One interval of a batched-ensemble run, read on the device by the captured
// step (scalars and the interval's record pair).
struct EnsStep {
    MoveParams P;
    RecordPair<float> F;
};

// Compact ensemble state, structure of arrays: cell = jI*xi + iI with FP32
// offsets fj, fi in [0, 1) and FP32 level k; FP32 health; FP64 age and
// emission time; the particle's scenario.
struct EnsSoA {
    uint64_t *id;
    uint32_t *cell;
    float *fj, *fi, *k, *health;
    double *age, *time;
    uint8_t *scen;
};

// Drift object type and crosswind side of a particle (particle_data's
// driftObjectType and driftSide) packed in one word: type in bits 0-15, the
// side's two's-complement byte in bits 16-23.  Type 0 is PASSIVE.
__host__ __device__ __forceinline__ uint32_t make_tag(uint16_t type, int8_t side) {
    return (uint32_t)type | ((uint32_t)(uint8_t)side << 16);
}
__host__ __device__ __forceinline__ uint16_t tag_type(uint32_t tag) { return (uint16_t)(tag & 0xffffu); }
__host__ __device__ __forceinline__ int8_t tag_side(uint32_t tag) { return (int8_t)(uint8_t)(tag >> 16); }

// Drift fields of released particles (Source.cpp:38-42): the configured type,
// and the configured side or, with randomSide, RIGHT when the keyed uniform
// of the particle is below rightProbability, else LEFT.
struct EmitDrift {
    uint16_t type;
    int8_t side;
    int randomSide;
    double rightProbability;
};

// Largest population of one launch: every kernel indexes particles with a
// 32-bit int (blockIdx.x * blockDim.x + threadIdx.x) and sizes its grid as
// (n + block - 1) / block, which both stay below 2^31 for n up to this.
constexpr int kMaxParticles = 0x7FFFFF00;

// Particle state, structure of arrays (the context's canonical layout).  ev
// holds each particle's closure events (WC_EVENT_* bits, OR-ed by the move
// kernels when event tracking is on); it travels through compaction.
struct SoA {
    uint64_t *id;
    double *k, *j, *i, *health, *age, *time;
    uint32_t *tag;
    uint8_t *ev;
};

struct State {
    uint64_t id;
    double k, j, i, health, age, time;
    uint32_t tag;
};

// The tag is read only when the move can change it (tags: the jibe is on),
// and written back for non-passive particles alone.
struct SoAAccess {
    SoA s;
    int tags;
    __device__ __forceinline__ void load(int idx, State &p) const {
        p.id = s.id[idx]; p.k = s.k[idx]; p.j = s.j[idx]; p.i = s.i[idx];
        p.health = s.health[idx]; p.age = s.age[idx]; p.time = s.time[idx];
        p.tag = tags ? s.tag[idx] : 0u;
    }
    __device__ __forceinline__ void store(int idx, const State &p) const {
        s.k[idx] = p.k; s.j[idx] = p.j; s.i[idx] = p.i; s.health[idx] = p.health; s.age[idx] = p.age;
        if (tags && tag_type(p.tag) != 0) s.tag[idx] = p.tag;
    }
};

struct AoSAccess {
    wc_particle_data *p;
    __device__ __forceinline__ void load(int idx, State &q) const {
        const wc_particle_data &d = p[idx];
        q.id = d.id; q.k = d.k; q.j = d.j; q.i = d.i; q.health = d.health; q.age = d.age; q.time = d.time;
        q.tag = make_tag(d.driftObjectType, d.driftSide);
    }
    __device__ __forceinline__ void store(int idx, const State &q) const {
        wc_particle_data &d = p[idx];
        d.k = q.k; d.j = q.j; d.i = q.i; d.health = q.health; d.age = q.age;
        if (tag_type(q.tag) != 0) d.driftSide = tag_side(q.tag);
    }
};

// SplitMix64 finaliser and ordered key combination, scheme
// ordered-splitmix64-v2 (NumericalHelpers.hpp:84-112).
__device__ __forceinline__ uint64_t mix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

// The hourly jibe of a non-passive particle after a live substep
// (Particle.cpp:805-813): true when the keyed uniform of component 'JIBE'
// at substep floor(elapsed / dti) falls below jibeStepProbability
// (DriftModel.hpp:156-160).  keyPrefix = the key after seed, id and
// interval.  FP64 with the reference's operation order on every path.
__device__ __forceinline__ bool jibe_flips(const MoveParams &P, uint64_t keyPrefix, double elapsed, double stepDt) {
#pragma clang fp contract(off)
    const double prob = stepDt <= 0 ? 0.0 : P.jibeHourly >= 1 ? 1.0 : 1 - exp(log(1 - P.jibeHourly) * stepDt / 3600);
    const uint64_t sk = mix64(keyPrefix ^ (uint64_t)floor(elapsed / P.dti));
    return (double)(mix64(sk ^ 0x4a494245ULL) >> 11) * 0x1.0p-53 < prob;
}

__device__ __forceinline__ uint32_t flip_side(uint32_t tag) {
    return make_tag(tag_type(tag), (int8_t)-(int)tag_side(tag));
}

// Sum of a 64-bit value over the wavefront, added once per wave to *counter.
__device__ __forceinline__ void wave_add_u64(unsigned long long *counter, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(counter, v);
}

}  // namespace wc
