// oe_noalloc.hip -- entry points for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Their kernels keep partial results in a file-scope
// device array instead of library- or caller-allocated scratch.
//
// The array's name is only a host-side stand-in for the symbol: host code
// uses the per-device address hipGetSymbolAddress returns for it, looked up
// once per device with that device current, and never the name itself.
#include "oe_objects.h"

#include <mutex>

namespace {

// Partial sums of the energy reduction: one (E, H) pair per block of the
// fixed partition the engines use, so results equal oe_engine_energy's.
constexpr size_t SCRATCH_DOUBLES = 2 * 1024;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on `device` (current), looked up once.
oe_status scratch_on(int device, double** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !p)
            return oe::fail(OE_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                            hipGetErrorString(rc));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return OE_OK;
}

}  // namespace

extern "C" {

oe_status oe_fields_energy_noalloc(const float* fields, size_t n, double* out, int device, void* stream) {
    if (!fields || !out) return oe::fail(OE_ERR_CONFIG, "null argument");
    if (n == 0) return oe::fail(OE_ERR_GRID, "empty field block");
    if (n > OE_NOALLOC_MAX_CELLS)
        return oe::fail(OE_ERR_GRID, "field block of %zu cells exceeds the no-allocation limit of %llu cells", n,
                        (unsigned long long)OE_NOALLOC_MAX_CELLS);
    if (device < 0 || device >= MAX_DEVICES) return oe::fail(OE_ERR_CONFIG, "device %d out of range", device);
    if ((size_t)2 * oe::energy_blocks() > SCRATCH_DOUBLES)
        return oe::fail(OE_ERR_DEVICE, "energy partition of %d blocks exceeds the scratch array", oe::energy_blocks());
    int prev = 0;
    HIP_OK(hipGetDevice(&prev));
    if (prev != device) HIP_OK(hipSetDevice(device));
    double* partials = nullptr;
    oe_status st = scratch_on(device, &partials);
    if (st == OE_OK) {
        // every block writes its own pair before the final pass reads it, so
        // nothing from an earlier call needs clearing
        const hipError_t rc = oe::launch_energy(fields, n, partials, oe::energy_blocks(), out, (hipStream_t)stream);
        if (rc != hipSuccess)
            st = oe::fail(OE_ERR_DEVICE, "energy launch failed: %s", hipGetErrorString(rc));
    }
    if (prev != device) HIP_OK(hipSetDevice(prev));
    return st;
}

}  // extern "C"
