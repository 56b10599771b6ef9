// oe_coef.hip -- the material table: the Yee updates read a 2-byte index per
// cell into a small table of update coefficients instead of six 4-byte
// coefficient streams per half-step.
//
// The reference stores Ca, Cb, Da, Db per cell and component (12 f32 arrays,
// src/fdtd/operator.rs:92-141), but a grid built from material boxes holds
// only as many distinct coefficient sets as it has materials: the operator
// uses one spacing per axis (the minimum, src/fdtd/operator.rs:82-89), so a
// cell's six H coefficients (Da, Db of x, y, z) and six E coefficients (Ca,
// Cb) depend only on what was written into it.  The table is derived from
// the coefficient arrays themselves, so it holds exactly their float values
// and the updates stay bit-identical:
//
//   claim   one pass over the cells: a cell whose tuple differs from the
//           previous cell's hashes its 6-float tuple (bit patterns) into a
//           4096-slot open-addressing table; the first cell with a new tuple
//           claims a free slot (compare-and-swap of its cell index), every
//           other cell compares its tuple exactly with the owner's.  More
//           than `cap` distinct tuples, or a probe sequence longer than 64
//           slots, sets the overflow flag and the engine keeps the 12-array
//           path.
//   number  the host numbers the claimed slots; a gather builds the table
//           rows from the owners' coefficients.
//   assign  a second pass writes every cell's table row (uint16).
//
// One half-step reads: H -- Da, Db of x, y, z (arrays 6..11); E -- Ca, Cb
// (arrays 0..5).  Table rows are laid out (a_x, b_x, a_y, b_y, a_z, b_z).

#include "oe_internal.h"

namespace oe {
namespace {

constexpr int SLOTS = 4096;
constexpr int MAX_PROBE = 64;
constexpr uint32_t EMPTY = 0xffffffffu;

struct Rows {
    const float* r[6];  // a_x, b_x, a_y, b_y, a_z, b_z arrays
};

__device__ __forceinline__ void tuple_of(const Rows& s, size_t c, uint32_t* w) {
#pragma unroll
    for (int q = 0; q < 6; ++q) w[q] = __float_as_uint(s.r[q][c]);
}

__device__ __forceinline__ uint32_t hash6(const uint32_t* w) {
    uint32_t h = 0x9e3779b9u;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        h ^= w[q] + 0x7f4a7c15u + (h << 6) + (h >> 2);
        h *= 0x85ebca6bu;
    }
    h ^= h >> 16;
    h *= 0xc2b2ae35u;
    h ^= h >> 13;
    return h;
}

__device__ __forceinline__ bool same(const Rows& s, uint32_t owner, const uint32_t* w) {
    bool eq = true;
#pragma unroll
    for (int q = 0; q < 6; ++q) eq = eq && (__float_as_uint(s.r[q][owner]) == w[q]);
    return eq;
}

__global__ void k_tab_claim(Rows s, size_t N, uint32_t* __restrict__ slot, int* __restrict__ count,
                            int* __restrict__ overflow, int cap) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += stride) {
        if (__hip_atomic_load(overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        uint32_t w[6];
        tuple_of(s, c, w);
        // The first cell of every distinct tuple differs from its predecessor,
        // so cells equal to the previous cell need not probe (runs along k
        // are long; this keeps millions of lanes off one hot slot).
        if (c > 0 && same(s, (uint32_t)(c - 1), w)) continue;
        uint32_t h = hash6(w) & (SLOTS - 1);
        bool done = false;
        for (int probe = 0; probe < MAX_PROBE && !done; ++probe) {
            uint32_t o = __hip_atomic_load(&slot[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o == EMPTY) {
                const uint32_t prev = atomicCAS(&slot[h], EMPTY, (uint32_t)c);
                if (prev == EMPTY) {
                    if (atomicAdd(count, 1) >= cap) atomicExch(overflow, 1);
                    done = true;
                    break;
                }
                o = prev;
            }
            if (same(s, o, w)) {
                done = true;
                break;
            }
            h = (h + 1) & (SLOTS - 1);
        }
        if (!done) atomicExch(overflow, 1);
    }
}

__global__ void k_tab_gather(Rows s, const uint32_t* __restrict__ slot, const int* __restrict__ slot_id,
                             float* __restrict__ tab) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= SLOTS) return;
    const uint32_t o = slot[h];
    if (o == EMPTY) return;
    const int id = slot_id[h];
#pragma unroll
    for (int q = 0; q < 6; ++q) tab[(size_t)id * 6 + q] = s.r[q][o];
}

__global__ void k_tab_assign(Rows s, size_t N, const uint32_t* __restrict__ slot, const int* __restrict__ slot_id,
                             uint16_t* __restrict__ mat, int* __restrict__ missing) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += stride) {
        uint32_t w[6];
        tuple_of(s, c, w);
        uint32_t h = hash6(w) & (SLOTS - 1);
        int id = -1;
        for (int probe = 0; probe < MAX_PROBE; ++probe) {
            const uint32_t o = slot[h];
            if (o == EMPTY) break;
            if (same(s, o, w)) {
                id = slot_id[h];
                break;
            }
            h = (h + 1) & (SLOTS - 1);
        }
        if (id < 0) {
            atomicExch(missing, 1);
            id = 0;
        }
        mat[c] = (uint16_t)id;
    }
}

Rows rows_of(const float* Cf, size_t N, bool h) {
    Rows s{};
    const int base = h ? 6 : 0;
    for (int d = 0; d < 3; ++d) {
        s.r[2 * d] = Cf + (size_t)(base + d) * N;
        s.r[2 * d + 1] = Cf + (size_t)(base + 3 + d) * N;
    }
    return s;
}

}  // namespace

// Builds one half's table into tab (cap x 6 floats) and mat (N entries).
// Returns the number of rows, or 0 when the grid holds more than cap
// distinct tuples (the caller then keeps the coefficient arrays).
hipError_t build_material_table(const float* Cf, size_t N, bool h, int cap, uint16_t* mat, float* tab,
                                int* n_rows, hipStream_t s) {
    *n_rows = 0;
    uint32_t* slot = nullptr;
    int* flags = nullptr;  // [count, overflow, missing]
    int* slot_id = nullptr;
    hipError_t err = hipSuccess;
    const Rows r = rows_of(Cf, N, h);
    std::vector<uint32_t> hslot(SLOTS);
    std::vector<int> hid(SLOTS, -1);
    int hf[3] = {0, 0, 0};
    int n = 0;
    const int grid = (int)std::min<size_t>((N + 255) / 256, 16384);
#define TRY(x) do { err = (x); if (err != hipSuccess) goto done; } while (0)
    TRY(hipMalloc(&slot, SLOTS * sizeof(uint32_t)));
    TRY(hipMalloc(&flags, 3 * sizeof(int)));
    TRY(hipMalloc(&slot_id, SLOTS * sizeof(int)));
    TRY(hipMemsetAsync(slot, 0xff, SLOTS * sizeof(uint32_t), s));
    TRY(hipMemsetAsync(flags, 0, 3 * sizeof(int), s));
    hipLaunchKernelGGL(k_tab_claim, dim3(grid), dim3(256), 0, s, r, N, slot, flags, flags + 1, cap);
    TRY(hipGetLastError());
    TRY(hipMemcpyAsync(hf, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    TRY(hipMemcpyAsync(hslot.data(), slot, SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    if (hf[1]) goto done;
    for (int q = 0; q < SLOTS; ++q)
        if (hslot[q] != EMPTY) hid[q] = n++;
    if (n > cap || n != hf[0]) goto done;
    TRY(hipMemcpyAsync(slot_id, hid.data(), SLOTS * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_tab_gather, dim3(SLOTS / 256), dim3(256), 0, s, r, slot, slot_id, tab);
    TRY(hipGetLastError());
    hipLaunchKernelGGL(k_tab_assign, dim3(grid), dim3(256), 0, s, r, N, slot, slot_id, mat, flags + 2);
    TRY(hipGetLastError());
    TRY(hipMemcpyAsync(hf, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    if (hf[2] == 0) *n_rows = n;
#undef TRY
done:
    if (slot) (void)hipFree(slot);
    if (flags) (void)hipFree(flags);
    if (slot_id) (void)hipFree(slot_id);
    return err;
}

namespace {
__global__ void k_narrow(const uint16_t* __restrict__ in, uint8_t* __restrict__ out, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        out[q] = (uint8_t)in[q];
}
}  // namespace

hipError_t narrow_material_index(const uint16_t* in, uint8_t* out, size_t n, hipStream_t s) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 16384);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_narrow, dim3(grid), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

}  // namespace oe
