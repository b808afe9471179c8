// Concentration gridding and dead-particle compaction for gfx950.
//
// Gridding (Concentration.hpp:16-44): each alive particle's (int)round(k, j, i)
// -- round half away from zero -- names a cell of the [s_rho][eta][xi] frame.
// Instead of one float atomic per particle (the reference's omp atomic), a
// workgroup of 1024 work-items counts 4096 consecutive particles into an LDS
// hash table of 8192 (cell, count) entries (linear probing, at most 16 probes,
// then a direct global add) and adds each distinct cell's count to the frame
// with one uint32 atomic.  Particles released together share few cells, so
// their counts meet in LDS instead of serialising on a few global addresses;
// dispersed particles cost one global atomic each.  Integer adds make the
// result exact and independent of scheduling.
//
// Compaction (Wacomm.cpp:710-713, erase(remove_if(!isAlive))): order-preserving
// stream compaction of the SoA state: block counts -> exclusive scan -> scatter
// with wave-ballot prefix sums inside each block.
#include "wc_device.h"
#include "wc_launch.h"

#pragma clang fp contract(off)

namespace wc {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Number of active lanes below this lane in a 64-bit ballot mask.
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
    return __popcll(mask & ((1ULL << lane_id()) - 1ULL));
}

constexpr int kGridBlock = 1024;                 // work-items per gridding workgroup
constexpr int kGridPerItem = 4;                  // particles per work-item
constexpr int kGridTableBits = 13;               // LDS table of 8192 entries (64 KiB)
constexpr uint32_t kGridEmpty = 0xFFFFFFFFu;     // no cell index reaches it (cells < 2^32 - 1)

// Out is uint32_t (exact counts, the context path) or float (the drop-in path,
// which adds into the caller's float frame like the reference does).
template <class Acc, class Out>
__global__ __launch_bounds__(kGridBlock) void grid_count_kernel(Acc acc, int n, int s_rho, int eta, int xi,
                                                                Out *__restrict__ counts) {
    constexpr uint32_t T = 1u << kGridTableBits;
    __shared__ uint32_t cellOf[T], countOf[T];
    for (uint32_t q = threadIdx.x; q < T; q += kGridBlock) {
        cellOf[q] = kGridEmpty;
        countOf[q] = 0;
    }
    __syncthreads();
    const long long base = (long long)blockIdx.x * (kGridBlock * kGridPerItem);
    for (int r = 0; r < kGridPerItem; ++r) {
        const long long idx = base + (long long)r * kGridBlock + threadIdx.x;
        if (idx >= n) break;
        double k, j, i, health;
        acc.position((int)idx, k, j, i, health);
        if (!(health > 0)) continue;                                      // Particle::isAlive
        const double kr = round(k), jr = round(j), ir = round(i);         // Concentration.hpp:31-33
        if (!(jr >= 0 && jr < eta && ir >= 0 && ir < xi && kr >= -s_rho + 1 && kr <= 0)) continue;   // :36
        const uint32_t key = (uint32_t)((((int)kr + s_rho - 1) * (size_t)eta + (int)jr) * xi + (int)ir);
        uint32_t slot = (key * 2654435761u) >> (32 - kGridTableBits);
        bool held = false;
        for (int probe = 0; probe < 16; ++probe) {
            const uint32_t old = atomicCAS(&cellOf[slot], kGridEmpty, key);
            if (old == kGridEmpty || old == key) {
                atomicAdd(&countOf[slot], 1u);
                held = true;
                break;
            }
            slot = (slot + 1) & (T - 1);
        }
        if (!held) atomicAdd(&counts[key], (Out)1);
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < T; q += kGridBlock)
        if (countOf[q]) atomicAdd(&counts[cellOf[q]], (Out)countOf[q]);
}

// perm (optional): visit the particles in this order (a cell-tile sort)
struct SoAPos {
    const double *k, *j, *i, *health;
    const uint32_t *perm;
    __device__ __forceinline__ void position(int t, double &K, double &J, double &I, double &H) const {
        const int idx = perm ? (int)perm[t] : t;
        K = k[idx]; J = j[idx]; I = i[idx]; H = health[idx];
    }
};
struct AoSPos {
    const wc_particle_data *p;
    __device__ __forceinline__ void position(int idx, double &K, double &J, double &I, double &H) const {
        K = p[idx].k; J = p[idx].j; I = p[idx].i; H = p[idx].health;
    }
};

// a[0..n) = v: 16-byte stores for the aligned body, 4-byte ones around it
__global__ void fill_u32_kernel(uint32_t *__restrict__ a, size_t n, uint32_t v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x, t0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t head = ((16 - ((uintptr_t)a & 15)) & 15) / 4;   // a is 4-byte aligned
    const size_t lead = head < n ? head : n;
    const size_t body = (n - lead) / 4;
    uint4 *b = reinterpret_cast<uint4 *>(a + lead);
    const uint4 vv = make_uint4(v, v, v, v);
    for (size_t q = t0; q < body; q += stride) b[q] = vv;
    for (size_t q = t0; q < lead; q += stride) a[q] = v;
    for (size_t q = lead + body * 4 + t0; q < n; q += stride) a[q] = v;
}

// counts -> the reference's float frame: +1.0f repeated `count` times from
// zero saturates at 2^24; land cells (mask != 1) get 1e37 when masking
// (Wacomm.cpp:724-739).  One work-item per horizontal cell walks the levels,
// so the mask is read once per cell and no index is divided.
__global__ void counts_to_conc_kernel(const uint32_t *__restrict__ counts, const double *__restrict__ mask,
                                      int mask_output, int s_rho, size_t plane, float *__restrict__ conc) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < plane; c += (size_t)gridDim.x * blockDim.x) {
        const bool land = mask_output && mask[c] != 1.0;
        for (int l = 0; l < s_rho; ++l) {
            const size_t q = (size_t)l * plane + c;
            const uint32_t n = counts[q];
            conc[q] = land ? 1e37f : (n > 16777216u ? 16777216.0f : (float)n);
        }
    }
}

// ---- compaction ----------------------------------------------------------

__global__ __launch_bounds__(kBlock) void alive_count_kernel(const double *__restrict__ health, int n,
                                                             uint32_t *__restrict__ blockCounts) {
    __shared__ uint32_t waveCount[kWaves];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool alive = idx < n && health[idx] > 0;
    const unsigned long long m = __ballot(alive);
    if (lane_id() == 0) waveCount[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kWaves; ++w) s += waveCount[w];
        blockCounts[blockIdx.x] = s;
    }
}

// Exclusive scan of nb block counts in place by one block of 1024 threads;
// total[0] receives the sum.
__global__ __launch_bounds__(1024) void scan_blocks_kernel(uint32_t *__restrict__ counts, int nb,
                                                           unsigned long long *__restrict__ total) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per;
    uint32_t s = 0;
    for (int q = 0; q < per && b0 + q < nb; ++q) s += counts[b0 + q];
    part[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partial sums
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t == 0 ? 0 : part[t - 1];
    for (int q = 0; q < per && b0 + q < nb; ++q) {
        const uint32_t c = counts[b0 + q];
        counts[b0 + q] = run;
        run += c;
    }
    if (t == 1023) total[0] = part[1023];
}

// remap (optional): remap[idx] = the survivor's new index, or kDead
constexpr uint32_t kDead = 0xFFFFFFFFu;
// Exclusive scan of n uint32 in place for large n, in three passes: the sum
// of every chunk of kScanChunk entries, a one-block scan of the sums, and
// the chunks' scans plus their offsets.  Each work-item takes 4 consecutive
// entries (one 16-byte load when aligned), a wave scans with shuffles.
constexpr int kScanChunk = kBlock * 4;

__device__ __forceinline__ void load4(const uint32_t *a, size_t n, size_t q, uint32_t v[4]) {
    for (int r = 0; r < 4; ++r) v[r] = q + r < n ? a[q + r] : 0u;
}

__global__ __launch_bounds__(kBlock) void chunk_sum_kernel(const uint32_t *__restrict__ a, size_t n,
                                                           uint32_t *__restrict__ sums) {
    __shared__ uint32_t waveSum[kWaves];
    uint32_t v[4];
    load4(a, n, (size_t)blockIdx.x * kScanChunk + 4 * threadIdx.x, v);
    uint32_t s = v[0] + v[1] + v[2] + v[3];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane_id() == 0) waveSum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kWaves; ++w) t += waveSum[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void chunk_scan_kernel(uint32_t *__restrict__ a, size_t n,
                                                            const uint32_t *__restrict__ offsets) {
    __shared__ uint32_t waveSum[kWaves];
    const size_t q = (size_t)blockIdx.x * kScanChunk + 4 * threadIdx.x;
    uint32_t v[4];
    load4(a, n, q, v);
    const uint32_t own = v[0] + v[1] + v[2] + v[3];
    uint32_t incl = own;                                   // inclusive scan over the wave
    const int lane = lane_id();
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) waveSum[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t base = offsets[blockIdx.x];
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) base += waveSum[w];
    uint32_t run = base + incl - own;
    for (int r = 0; r < 4; ++r)
        if (q + r < n) {
            const uint32_t x = v[r];
            a[q + r] = run;
            run += x;
        }
}

__global__ __launch_bounds__(kBlock) void scatter_alive_kernel(SoA src, SoA dst, int n,
                                                               const uint32_t *__restrict__ blockOffsets,
                                                               uint32_t *__restrict__ remap) {
    __shared__ uint32_t waveBase[kWaves];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool alive = idx < n && src.health[idx] > 0;
    const unsigned long long m = __ballot(alive);
    if (lane_id() == 0) waveBase[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = blockOffsets[blockIdx.x];
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) base += waveBase[w];
    if (alive) {
        const uint32_t o = base + (uint32_t)lanes_below(m);
        dst.id[o] = src.id[idx]; dst.k[o] = src.k[idx]; dst.j[o] = src.j[idx]; dst.i[o] = src.i[idx];
        dst.health[o] = src.health[idx]; dst.age[o] = src.age[idx]; dst.time[o] = src.time[idx];
        dst.tag[o] = src.tag[idx]; dst.ev[o] = src.ev[idx];
        if (remap) remap[idx] = o;
    } else if (remap && idx < n) {
        remap[idx] = kDead;
    }
}

// ---- cell-tile order of the population (a permutation) -----------------
//
// perm lists the population's indices with particles of one tile of
// tile x tile cells together (tiles row-major; particles outside the grid
// last), so the move's lanes gather from neighbouring cells.  A counting
// sort: histogram of tile keys, exclusive scan, scatter.  Both passes read
// the particles in the previous order, where a workgroup's particles share
// few tiles.

__device__ __forceinline__ uint32_t tile_of(double j, double i, int eta, int xi, int tile, int tilesX,
                                            uint32_t nTiles) {
    const double jf = floor(j), if_ = floor(i);
    if (!(jf >= 0 && if_ >= 0 && jf < eta && if_ < xi)) return nTiles;
    return (uint32_t)((int)jf / tile) * (uint32_t)tilesX + (uint32_t)((int)if_ / tile);
}

// Histogram and scatter of the tile keys, per workgroup of kGridBlock work-
// items x kGridPerItem particles, aggregated in an LDS hash table (as the
// gridding): one global atomic per distinct tile per workgroup.  The scatter
// reserves each tile's run of positions with one returning atomic and places
// the workgroup's particles of the tile in it.
struct TileTable {
    uint32_t key[1u << kGridTableBits], count[1u << kGridTableBits], base[1u << kGridTableBits];
};

__device__ __forceinline__ void table_clear(TileTable &T) {
    for (uint32_t q = threadIdx.x; q < (1u << kGridTableBits); q += kGridBlock) {
        T.key[q] = kGridEmpty;
        T.count[q] = 0;
    }
}

// slot of key (inserted if new), or -1 when 16 probes find none
__device__ __forceinline__ int table_slot(TileTable &T, uint32_t key) {
    uint32_t slot = (key * 2654435761u) >> (32 - kGridTableBits);
    for (int probe = 0; probe < 16; ++probe) {
        const uint32_t old = atomicCAS(&T.key[slot], kGridEmpty, key);
        if (old == kGridEmpty || old == key) return (int)slot;
        slot = (slot + 1) & ((1u << kGridTableBits) - 1);
    }
    return -1;
}

__global__ __launch_bounds__(kGridBlock) void sort_hist_kernel(const double *__restrict__ j,
                                                               const double *__restrict__ i,
                                                               const uint32_t *__restrict__ perm, int n, int eta,
                                                               int xi, int tile, int tilesX, uint32_t nTiles,
                                                               uint32_t *__restrict__ bins) {
    __shared__ TileTable T;
    table_clear(T);
    __syncthreads();
    const long long base = (long long)blockIdx.x * (kGridBlock * kGridPerItem);
    for (int r = 0; r < kGridPerItem; ++r) {
        const long long t = base + (long long)r * kGridBlock + threadIdx.x;
        if (t >= n) break;
        const int idx = perm ? (int)perm[t] : (int)t;
        const uint32_t key = tile_of(j[idx], i[idx], eta, xi, tile, tilesX, nTiles);
        const int slot = table_slot(T, key);
        if (slot >= 0) atomicAdd(&T.count[slot], 1u);
        else atomicAdd(&bins[key], 1u);
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < (1u << kGridTableBits); q += kGridBlock)
        if (T.count[q]) atomicAdd(&bins[T.key[q]], T.count[q]);
}

__global__ __launch_bounds__(kGridBlock) void sort_scatter_kernel(const double *__restrict__ j,
                                                                  const double *__restrict__ i,
                                                                  const uint32_t *__restrict__ perm, int n, int eta,
                                                                  int xi, int tile, int tilesX, uint32_t nTiles,
                                                                  uint32_t *__restrict__ bins,
                                                                  uint32_t *__restrict__ out) {
    __shared__ TileTable T;
    table_clear(T);
    __syncthreads();
    const long long base = (long long)blockIdx.x * (kGridBlock * kGridPerItem);
    int idxs[kGridPerItem], slots[kGridPerItem];
    uint32_t local[kGridPerItem];
    for (int r = 0; r < kGridPerItem; ++r) {
        const long long t = base + (long long)r * kGridBlock + threadIdx.x;
        idxs[r] = -1;
        if (t >= n) continue;
        const int idx = perm ? (int)perm[t] : (int)t;
        const uint32_t key = tile_of(j[idx], i[idx], eta, xi, tile, tilesX, nTiles);
        const int slot = table_slot(T, key);
        idxs[r] = idx;
        slots[r] = slot;
        if (slot >= 0) local[r] = atomicAdd(&T.count[slot], 1u);
        else out[atomicAdd(&bins[key], 1u)] = (uint32_t)idx;       // table full: placed directly
        if (slot < 0) idxs[r] = -1;
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < (1u << kGridTableBits); q += kGridBlock)
        if (T.count[q]) T.base[q] = atomicAdd(&bins[T.key[q]], T.count[q]);
    __syncthreads();
    for (int r = 0; r < kGridPerItem; ++r)
        if (idxs[r] >= 0) out[T.base[slots[r]] + local[r]] = (uint32_t)idxs[r];
}

// perm after a compaction: the surviving entries, in perm order, mapped to
// the survivors' new indices (remap from scatter_alive_kernel)
__global__ __launch_bounds__(kBlock) void perm_count_kernel(const uint32_t *__restrict__ perm,
                                                            const uint32_t *__restrict__ remap, int n,
                                                            uint32_t *__restrict__ blockCounts) {
    __shared__ uint32_t waveCount[kWaves];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool keep = t < n && remap[perm[t]] != kDead;
    const unsigned long long m = __ballot(keep);
    if (lane_id() == 0) waveCount[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kWaves; ++w) s += waveCount[w];
        blockCounts[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kBlock) void perm_scatter_kernel(const uint32_t *__restrict__ perm,
                                                              const uint32_t *__restrict__ remap, int n,
                                                              const uint32_t *__restrict__ blockOffsets,
                                                              uint32_t *__restrict__ out) {
    __shared__ uint32_t waveBase[kWaves];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t to = t < n ? remap[perm[t]] : kDead;
    const bool keep = to != kDead;
    const unsigned long long m = __ballot(keep);
    if (lane_id() == 0) waveBase[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = blockOffsets[blockIdx.x];
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) base += waveBase[w];
    if (keep) out[base + (uint32_t)lanes_below(m)] = to;
}

__global__ void perm_fill_kernel(uint32_t *__restrict__ perm, size_t from, size_t to) {
    for (size_t q = from + blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < to; q += (size_t)gridDim.x * blockDim.x)
        perm[q] = (uint32_t)q;
}

// ---- particle history records ------------------------------------------

// Particles::netcdfRecords (Particles.cpp:115-147) with
// OceanModelAdapter::kji2deplatlon (OceanModelAdapter.cpp:271-321): a particle
// is kept when j >= 0, i >= 0, k <= 0, its cell (floor j, floor i) and its
// next cell lie on the grid, ceil(k) lies in (-s_w+1, 0], and none of depth,
// lat, lon is the 1e37 sentinel.  The range tests compare in double, so NaN or
// out-of-int-range coordinates are rejected as the host's int conversion
// (INT_MIN) rejects them.  The bilinear sums keep the reference's term order.
__device__ __forceinline__ bool record_geo(const GeoPtrs &g, double k, double j, double i, double &dep, double &lat,
                                           double &lon) {
    if (j < 0 || i < 0 || k > 0) return false;
    const double kc = ceil(k), jf = floor(j), fi = floor(i);
    if (!(jf >= 0.0 && fi >= 0.0 && jf + 1.0 < (double)g.eta && fi + 1.0 < (double)g.xi && kc <= 0.0 &&
          kc > (double)(-g.s_w + 1)))
        return false;
    const int jI = (int)jf, iI = (int)fi, kI = (int)kc;
    const double kF = kc - k, jF = j - jf, iF = i - fi;
    const size_t c00 = (size_t)jI * g.xi + iI, c10 = c00 + g.xi, c11 = c10 + 1, c01 = c00 + 1;
    lon = g.lon[c00] * (1.0 - iF) * (1.0 - jF) + g.lon[c10] * (1.0 - iF) * jF + g.lon[c11] * iF * jF +
          g.lon[c01] * iF * (1.0 - jF);
    lat = g.lat[c00] * (1.0 - iF) * (1.0 - jF) + g.lat[c10] * (1.0 - iF) * jF + g.lat[c11] * iF * jF +
          g.lat[c01] * iF * (1.0 - jF);
    const double h = g.h[c00] * (1.0 - iF) * (1.0 - jF) + g.h[c10] * (1.0 - iF) * jF + g.h[c11] * iF * jF +
                     g.h[c01] * iF * (1.0 - jF);
    const double sLo = g.sW[kI - 1 + g.s_w - 1], sUp = g.sW[kI + g.s_w - 1];
    const double aDep = fabs(fabs(h * sLo) - fabs(h * sUp));
    dep = -(h * fabs(sUp) + fabs(kF * aDep));
    return !(dep == 1e37 || lat == 1e37 || lon == 1e37);
}

__global__ __launch_bounds__(kBlock) void records_count_kernel(SoA src, int n, GeoPtrs g,
                                                               uint32_t *__restrict__ blockCounts) {
    __shared__ uint32_t waveCount[kWaves];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    double d, la, lo;
    const bool keep = idx < n && record_geo(g, src.k[idx], src.j[idx], src.i[idx], d, la, lo);
    const unsigned long long m = __ballot(keep);
    if (lane_id() == 0) waveCount[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kWaves; ++w) s += waveCount[w];
        blockCounts[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kBlock) void records_scatter_kernel(SoA src, int n, GeoPtrs g, SoA dst,
                                                                 double *__restrict__ dep, double *__restrict__ lat,
                                                                 double *__restrict__ lon,
                                                                 const uint32_t *__restrict__ blockOffsets) {
    __shared__ uint32_t waveBase[kWaves];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0, la = 0, lo = 0;
    const bool keep = idx < n && record_geo(g, src.k[idx], src.j[idx], src.i[idx], d, la, lo);
    const unsigned long long m = __ballot(keep);
    if (lane_id() == 0) waveBase[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = blockOffsets[blockIdx.x];
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) base += waveBase[w];
    if (keep) {
        const uint32_t o = base + (uint32_t)lanes_below(m);
        dst.id[o] = src.id[idx]; dst.k[o] = src.k[idx]; dst.j[o] = src.j[idx]; dst.i[o] = src.i[idx];
        dst.health[o] = src.health[idx]; dst.age[o] = src.age[idx]; dst.time[o] = src.time[idx];
        dst.tag[o] = src.tag[idx];
        dep[o] = d; lat[o] = la; lon[o] = lo;
    }
}

// Exclusive scan of a[0..n) in place; total[0] = the sum.  sums: scratch of
// kScanSumsCap entries (nullptr, or too few: the one-block scan).
void scan_u32(uint32_t *a, size_t n, uint32_t *sums, unsigned long long *total, hipStream_t stream) {
    const size_t chunks = (n + kScanChunk - 1) / kScanChunk;
    if (!sums || chunks <= 1 || chunks > kScanSumsCap) {
        scan_blocks_kernel<<<1, 1024, 0, stream>>>(a, (int)n, total);
        return;
    }
    chunk_sum_kernel<<<(int)chunks, kBlock, 0, stream>>>(a, n, sums);
    scan_blocks_kernel<<<1, 1024, 0, stream>>>(sums, (int)chunks, total);
    chunk_scan_kernel<<<(int)chunks, kBlock, 0, stream>>>(a, n, sums);
}

int grid_count_blocks(int n) {
    const long long per = (long long)kGridBlock * kGridPerItem;
    return (int)((n + per - 1) / per);
}

int grid_blocks(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    if (b > 65536) b = 65536;
    return (int)(b ? b : 1);
}

}  // namespace

hipError_t launch_fill_u32(uint32_t *a, size_t n, uint32_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    fill_u32_kernel<<<grid_blocks(n / 4 + 1), kBlock, 0, stream>>>(a, n, v);
    return hipGetLastError();
}

hipError_t launch_grid_count_soa(const SoA &s, int n, int s_rho, int eta, int xi, uint32_t *counts,
                                 const uint32_t *perm, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    grid_count_kernel<SoAPos, uint32_t><<<grid_count_blocks(n), kGridBlock, 0, stream>>>(
        SoAPos{s.k, s.j, s.i, s.health, perm}, n, s_rho, eta, xi, counts);
    return hipGetLastError();
}

hipError_t launch_grid_count_aos(const wc_particle_data *p, int n, int s_rho, int eta, int xi, uint32_t *counts,
                                 hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    grid_count_kernel<AoSPos, uint32_t><<<grid_count_blocks(n), kGridBlock, 0, stream>>>(AoSPos{p}, n, s_rho, eta,
                                                                                       xi, counts);
    return hipGetLastError();
}

// Adds each cell's count of this call into a float frame: exact while every
// cell stays at or below 2^24.
hipError_t launch_grid_count_aos_float(const wc_particle_data *p, int n, int s_rho, int eta, int xi, float *conc,
                                       hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    grid_count_kernel<AoSPos, float><<<grid_count_blocks(n), kGridBlock, 0, stream>>>(AoSPos{p}, n, s_rho, eta,
                                                                                    xi, conc);
    return hipGetLastError();
}

hipError_t launch_counts_to_conc(const uint32_t *counts, const double *mask, int mask_output, int s_rho,
                                 size_t plane, float *conc, hipStream_t stream) {
    counts_to_conc_kernel<<<grid_blocks(plane), kBlock, 0, stream>>>(counts, mask, mask_output, s_rho, plane,
                                                                      conc);
    return hipGetLastError();
}

// blockCounts must hold ceil(n/256) entries; total a device u64.
hipError_t launch_compact(const SoA &src, const SoA &dst, int n, uint32_t *blockCounts,
                          unsigned long long *total, uint32_t *remap, uint32_t *scanSums, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int nb = (n + kBlock - 1) / kBlock;
    alive_count_kernel<<<nb, kBlock, 0, stream>>>(src.health, n, blockCounts);
    scan_u32(blockCounts, (size_t)nb, scanSums, total, stream);
    scatter_alive_kernel<<<nb, kBlock, 0, stream>>>(src, dst, n, blockCounts, remap);
    return hipGetLastError();
}

hipError_t launch_sort_perm(const SoA &s, int n, int eta, int xi, int tile, const uint32_t *permIn,
                            uint32_t *permOut, uint32_t *bins, unsigned long long *total, uint32_t *scanSums,
                            hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int tilesX = (xi + tile - 1) / tile, tilesY = (eta + tile - 1) / tile;
    const uint32_t nTiles = (uint32_t)tilesX * (uint32_t)tilesY, nBins = nTiles + 1;
    const int nb = grid_count_blocks(n);
    fill_u32_kernel<<<grid_blocks(nBins / 4 + 1), kBlock, 0, stream>>>(bins, nBins, 0u);
    sort_hist_kernel<<<nb, kGridBlock, 0, stream>>>(s.j, s.i, permIn, n, eta, xi, tile, tilesX, nTiles, bins);
    scan_u32(bins, nBins, scanSums, total, stream);
    sort_scatter_kernel<<<nb, kGridBlock, 0, stream>>>(s.j, s.i, permIn, n, eta, xi, tile, tilesX, nTiles, bins,
                                                       permOut);
    return hipGetLastError();
}

hipError_t launch_perm_remap(const uint32_t *perm, const uint32_t *remap, int n, uint32_t *out,
                             uint32_t *blockCounts, unsigned long long *total, uint32_t *scanSums,
                             hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int nb = (n + kBlock - 1) / kBlock;
    perm_count_kernel<<<nb, kBlock, 0, stream>>>(perm, remap, n, blockCounts);
    scan_u32(blockCounts, (size_t)nb, scanSums, total, stream);
    perm_scatter_kernel<<<nb, kBlock, 0, stream>>>(perm, remap, n, blockCounts, out);
    return hipGetLastError();
}

hipError_t launch_perm_fill(uint32_t *perm, size_t from, size_t to, hipStream_t stream) {
    if (to <= from) return hipSuccess;
    perm_fill_kernel<<<grid_blocks(to - from), kBlock, 0, stream>>>(perm, from, to);
    return hipGetLastError();
}

}  // namespace wc

namespace wc {

// blockCounts must hold ceil(n/256) entries; total a device u64.
hipError_t launch_particle_records(const SoA &src, int n, const GeoPtrs &g, const SoA &dst, double *dep,
                                   double *lat, double *lon, uint32_t *blockCounts, unsigned long long *total,
                                   hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int nb = (n + kBlock - 1) / kBlock;
    records_count_kernel<<<nb, kBlock, 0, stream>>>(src, n, g, blockCounts);
    scan_blocks_kernel<<<1, 1024, 0, stream>>>(blockCounts, nb, total);
    records_scatter_kernel<<<nb, kBlock, 0, stream>>>(src, n, g, dst, dep, lat, lon, blockCounts);
    return hipGetLastError();
}

}  // namespace wc

