// C-ABI shim over reference headers compiled in place from the reference tree
// (include path set by oracle/Makefile).  Test infrastructure only: it lets the
// CPU tests compare the oracle restatement with the reference's own
// NumericalHelpers.hpp, SourceEmission.hpp and the inline helpers of
// DriftModel.hpp bit for bit.
#include "DriftModel.hpp"
#include "NumericalHelpers.hpp"
#include "SourceEmission.hpp"

#include <cstdint>
#include <cstring>
#include <stdexcept>

extern "C" {

std::uint64_t ref_mix(std::uint64_t v) { return NumericalHelpers::mix(v); }

double ref_uniform(std::uint64_t seed, std::uint64_t particle, std::int64_t interval, std::uint64_t substep,
                   std::uint64_t component) {
    return NumericalHelpers::uniform(seed, particle, interval, substep, component);
}

double ref_normal(std::uint64_t seed, std::uint64_t particle, std::int64_t interval, std::uint64_t substep,
                  std::uint64_t component) {
    return NumericalHelpers::normal(seed, particle, interval, substep, component);
}

double ref_restart_elapsed(double s, double e, double c) { return NumericalHelpers::restartElapsed(s, e, c); }

double ref_reflect_cell(double oldCoordinate, int oldCell, double candidate, int candidateCell) {
    NumericalHelpers::reflectCell(oldCoordinate, oldCell, candidate, candidateCell);
    return candidate;
}

double ref_reflect_domain(double coordinate, double maximum) {
    return NumericalHelpers::reflectDomain(coordinate, maximum);
}

const char *ref_scheme() { return NumericalHelpers::RANDOM_KEY_SCHEME; }

double ref_jibe_step_probability(double hourlyProbability, double elapsedSeconds) {
    return jibeStepProbability(hourlyProbability, elapsedSeconds);
}

int ref_sample_drift_side(std::uint64_t seed, std::uint64_t particle, double rightProbability) {
    return static_cast<int>(sampleDriftSide(seed, particle, rightProbability));
}

std::uint64_t ref_jibe_component() { return LEEWAY_JIBE_RANDOM_COMPONENT; }

std::uint64_t ref_partition_count(std::uint64_t total, std::uint64_t workers, std::uint64_t worker) {
    return NumericalHelpers::partitionCount(total, workers, worker);
}

std::uint64_t ref_partition_offset(std::uint64_t total, std::uint64_t workers, std::uint64_t worker) {
    return NumericalHelpers::partitionOffset(total, workers, worker);
}

// Same return convention as wco_emission_times: count, -1 invalid (throws), -2 cap too small.
long ref_emission_times(int mode, double value, double sourceStart, double sourceEnd, double intervalStart,
                        double intervalEnd, double *times, long cap) {
    EmissionMode m = mode == 0 ? EmissionMode::UNIFORM_RATE
                   : mode == 1 ? EmissionMode::FORCING_INTERVAL_BATCH : EmissionMode::SINGLE_PULSE;
    try {
        std::vector<double> t = sourceEmissionTimes(m, value, sourceStart, sourceEnd, intervalStart, intervalEnd);
        if ((long)t.size() > cap) return -2;
        if (!t.empty()) std::memcpy(times, t.data(), t.size() * sizeof(double));
        return (long)t.size();
    } catch (const std::exception &) {
        return -1;
    }
}

double ref_source_position_offset(std::uint64_t seed, std::uint64_t particle, std::uint64_t component) {
    return sourcePositionOffset(seed, particle, component);
}

}
