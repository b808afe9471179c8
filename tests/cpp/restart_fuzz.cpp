// Corrupt-input check of the restart reader (wc_read_restart_nc), built on
// the CPU with AddressSanitizer and UndefinedBehaviorSanitizer around the
// reader's own source: byte flips, truncations and header fields set to huge
// or random values must give a clean status (or a population that fits the
// file), never an out-of-bounds access.
//
//   restart_fuzz <valid restart file> <cases> <seed> <scratch file>
#include "wacomm_hip.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static uint64_t state;
static uint64_t next() {  // SplitMix64 stream for the mutations
    uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

int main(int argc, char **argv) {
    if (argc != 5) return 2;
    FILE *fp = fopen(argv[1], "rb");
    if (!fp) return 2;
    std::vector<uint8_t> good;
    int c;
    while ((c = fgetc(fp)) != EOF) good.push_back((uint8_t)c);
    fclose(fp);
    const long cases = atol(argv[2]);
    state = strtoull(argv[3], nullptr, 10);
    wc_config_data cfg{};
    cfg.randomSeed = 5489;
    cfg.trackingDirection = 1;
    long accepted = 0, rejected = 0;
    const size_t span = good.size() < 1400 ? good.size() : 1400;
    for (long k = 0; k < cases; ++k) {
        std::vector<uint8_t> b = good;
        switch (k % 4) {
        case 0:
            for (uint64_t q = 0, m = 1 + next() % 4; q < m; ++q) b[4 + next() % (span - 4)] ^= (uint8_t)(1u << (next() % 8));
            break;
        case 1: b.resize(next() % b.size()); break;
        case 2: {
            const size_t at = 4 * (1 + next() % (span / 4 - 1));
            const uint32_t v = (next() & 1) ? 0xFFFFFFFFu : (uint32_t)next();
            for (int q = 0; q < 4; ++q) b[at + q] = (uint8_t)(v >> (24 - 8 * q));
            break;
        }
        default: {
            const size_t at = 8 * (1 + next() % (span / 8 - 1));
            const uint64_t pick = next() % 3;
            const uint64_t v = pick == 0 ? 0x7FFFFFFFFFFFFFFFULL : pick == 1 ? 0xFFFFFFFFFFFFFFF8ULL : next();
            for (int q = 0; q < 8; ++q) b[at + q] = (uint8_t)(v >> (56 - 8 * q));
        }
        }
        FILE *out = fopen(argv[4], "wb");
        if (!out) return 2;
        if (!b.empty()) fwrite(b.data(), 1, b.size(), out);
        fclose(out);
        size_t n = 0;
        double cp = 0;
        char err[256];
        int rc = wc_read_restart_nc(argv[4], &cfg, "roms", nullptr, 0, &n, &cp, err, sizeof err);
        if (rc == WC_OK) {
            if (n * 8 > b.size()) {
                fprintf(stderr, "case %ld: %zu particles from %zu bytes\n", k, n, b.size());
                return 1;
            }
            std::vector<wc_particle_data> p(n ? n : 1);
            rc = wc_read_restart_nc(argv[4], &cfg, "roms", p.data(), n, &n, &cp, err, sizeof err);
        }
        if (rc == WC_OK) ++accepted;
        else ++rejected;
    }
    printf("ok %ld %ld\n", accepted, rejected);
    return 0;
}
