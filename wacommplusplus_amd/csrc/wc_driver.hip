// Interval driver and particle release on the device.
//
// wc_emit restates Sources::emit (Sources.cpp:233-248), Source::emit
// (Source.cpp:25-44) and sourceEmissionTimes (SourceEmission.hpp:38-79): the
// release times of every source are computed on the host (a few per source
// and interval), merged by a stable sort on time, and the new particles are
// written straight into the device SoA by one kernel -- ids continue from the
// last particle's id, and with randomSources each coordinate gets the keyed
// offset 0.25*normal(seed, id, SOURCE_POSITION_RANDOM_INTERVAL, 0, c).
//
// wc_run_window restates the interval loop of Wacomm::run (Wacomm.cpp:200-770)
// on one device: restart clipping, emission at the interval start, the move,
// order-preserving dead removal and gridding into the target frame, with the
// forcing pair held in a two-slot record ring (one upload per interval).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "wacomm_hip.h"
#include "wc_device.h"

#pragma clang fp contract(off)

namespace wc {
namespace {

constexpr int64_t kSourcePositionInterval = (-9223372036854775807LL - 1) + 2;   // SourceEmission.hpp:18
constexpr int64_t kLeewaySideInterval = (-9223372036854775807LL - 1) + 1;       // DriftModel.hpp:110-111

__device__ double keyed_normal_full(uint64_t seed, uint64_t particle, int64_t interval, uint64_t substep,
                                    uint64_t component) {
    // NumericalHelpers::normal (NumericalHelpers.hpp:104-120)
    const uint64_t k = mix64(mix64(mix64(mix64(seed) ^ particle) ^ (uint64_t)interval) ^ substep);
    double u1 = (double)(mix64(k ^ (component * 2)) >> 11) * 0x1.0p-53;
    u1 = (u1 < 2.2250738585072014e-308) ? 2.2250738585072014e-308 : u1;
    const double u2 = (double)(mix64(k ^ (component * 2 + 1)) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.28318530717958647692 * u2);
}

__global__ void emit_kernel(SoA s, size_t at, int n, const double *__restrict__ times, const int *__restrict__ src,
                            const double *__restrict__ pos, uint64_t firstId, uint64_t seed, int randomSources,
                            EmitDrift drift) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint64_t id = firstId + (uint64_t)q;
    const int sidx = src[q];
    double k = pos[3 * sidx], j = pos[3 * sidx + 1], i = pos[3 * sidx + 2];
    if (randomSources) {                                      // Source.cpp:31-35
        k = k + 0.25 * keyed_normal_full(seed, id, kSourcePositionInterval, 0, 0);
        j = j + 0.25 * keyed_normal_full(seed, id, kSourcePositionInterval, 0, 1);
        i = i + 0.25 * keyed_normal_full(seed, id, kSourcePositionInterval, 0, 2);
    }
    const size_t o = at + q;
    s.id[o] = id;
    s.k[o] = k;
    s.j[o] = j;
    s.i[o] = i;
    s.health[o] = 1.0;                                        // Particle(id,k,j,i,time): health0, age 0
    s.age[o] = 0.0;
    s.time[o] = times[q];
    int8_t side = drift.side;
    if (drift.randomSide) {                                   // sampleDriftSide, DriftModel.hpp:146-154
        const uint64_t key = mix64(mix64(mix64(mix64(mix64(seed) ^ id) ^ (uint64_t)kLeewaySideInterval) ^ 0ULL) ^ 0ULL);
        side = (double)(key >> 11) * 0x1.0p-53 < drift.rightProbability ? 1 : -1;
    }
    s.tag[o] = make_tag(drift.type, side);
    s.ev[o] = 0;
}

}  // namespace

hipError_t launch_emit(const SoA &s, size_t at, int n, const double *times, const int *src, const double *pos,
                       uint64_t firstId, uint64_t seed, int randomSources, EmitDrift drift, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    emit_kernel<<<(n + 255) / 256, 256, 0, stream>>>(s, at, n, times, src, pos, firstId, seed, randomSources,
                                                      drift);
    return hipGetLastError();
}

// sourceEmissionTimes (SourceEmission.hpp:38-79); returns false for an
// invalid schedule (the reference throws).
bool emission_times(int mode, double value, double sourceStart, double sourceEnd, double intervalStart,
                    double intervalEnd, std::vector<double> &out) {
    out.clear();
    if (!std::isfinite(value) || value < 0 || !std::isfinite(sourceStart) || !std::isfinite(intervalStart) ||
        !std::isfinite(intervalEnd) || intervalEnd <= intervalStart)
        return false;
    if (value == 0) return true;
    if (mode == 2) {                                          // SINGLE_PULSE
        if (std::floor(value) != value) return false;
        if (sourceStart >= intervalStart && sourceStart < intervalEnd) out.assign((size_t)value, sourceStart);
        return true;
    }
    if (mode == 1) {                                          // FORCING_INTERVAL_BATCH
        if (std::floor(value) != value) return false;
        const long long record = std::llround(intervalStart);
        if (record >= std::llround(sourceStart) && (sourceEnd < 0 || record <= std::llround(sourceEnd)))
            out.assign((size_t)value, intervalStart);
        return true;
    }
    const double activeStart = std::max(intervalStart, sourceStart);   // UNIFORM_RATE
    const double activeEnd = sourceEnd < 0 ? intervalEnd : std::min(intervalEnd, sourceEnd);
    if (activeEnd <= activeStart) return true;
    const double spacing = 3600.0 / value;
    const long long first = std::max(0LL, (long long)std::ceil((activeStart - sourceStart) / spacing - 0.5));
    for (long long index = first;; index++) {
        const double time = sourceStart + (index + 0.5) * spacing;
        if (time >= activeEnd) break;
        if (time >= activeStart) out.push_back(time);
    }
    return true;
}

}  // namespace wc
