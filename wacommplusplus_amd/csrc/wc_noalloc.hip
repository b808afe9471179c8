// Allocation-free gridding (wc_grid_noalloc), for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops.
//
// The exact per-cell counts accumulate in `scratch`, a file-scope device
// array of WC_GRID_NOALLOC_MAX_CELLS uint32 (64 MiB) that every device the
// library runs on holds until the process ends.  On the host the array's name
// is only a stand-in for the symbol, never a device address: the address comes
// from hipGetSymbolAddress, looked up once per device with that device current
// (at the first call on it) and kept in scratchPtr[device]; the kernels get
// that pointer as an argument.  All calls on one device share the array, so
// calls that overlap on a device race; each call clears the cells it uses
// first (with a fill kernel, so the sequence can be captured in a graph).
#include <hip/hip_runtime.h>

#include <mutex>

#include "wacomm_hip.h"
#include "wc_launch.h"

namespace wc {
namespace {

constexpr size_t kScratchCells = WC_GRID_NOALLOC_MAX_CELLS;
constexpr int kScratchDevices = 64;

__device__ __align__(256) uint32_t scratch[kScratchCells];

std::mutex scratchLock;
uint32_t *scratchPtr[kScratchDevices] = {};

// Device address of the current device's scratch array.
hipError_t scratch_address(int device, uint32_t **out) {
    std::lock_guard<std::mutex> guard(scratchLock);
    if (!scratchPtr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratchPtr[device] = static_cast<uint32_t *>(p);
    }
    *out = scratchPtr[device];
    return hipSuccess;
}

}  // namespace
}  // namespace wc

extern "C" int wc_grid_noalloc(const wc_particle_data *particles, int particleCount, int s_rho, int eta, int xi,
                               const double *mask, int mask_output, float *conc, void *stream) {
    using namespace wc;
    if (particleCount < 0 || particleCount > kMaxParticles || s_rho < 1 || eta < 1 || xi < 1 || !conc ||
        (particleCount > 0 && !particles) || (mask_output && !mask))
        return WC_ERR_INVALID;
    const size_t plane = (size_t)eta * xi;
    if (plane > kScratchCells || plane * (size_t)s_rho > kScratchCells) return WC_ERR_INVALID;
    const size_t cells = plane * (size_t)s_rho;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return WC_ERR_HIP;
    if (device < 0 || device >= kScratchDevices) return WC_ERR_INVALID;
    uint32_t *counts = nullptr;
    if (scratch_address(device, &counts) != hipSuccess) return WC_ERR_HIP;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if (launch_fill_u32(counts, cells, 0u, st) != hipSuccess) return WC_ERR_HIP;
    if (launch_grid_count_aos(particles, particleCount, s_rho, eta, xi, counts, st) != hipSuccess) return WC_ERR_HIP;
    if (launch_counts_to_conc(counts, mask, mask_output, s_rho, plane, conc, st) != hipSuccess) return WC_ERR_HIP;
    return WC_OK;
}
