// oe_objects.h -- the host objects behind the opaque handles of
// include/oe/oe.h (operator, engine, extension) and the error helpers, shared
// by the runtime files (oe_api.hip, oe_sweep.hip).  Not part of the public ABI.
#pragma once

#include "oe/oe.h"
#include "oe_internal.h"

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

namespace oe {
// Records a formatted message for oe_last_error and returns code.
oe_status fail(oe_status code, const char* fmt, ...);
}  // namespace oe

#define HIP_OK(expr)                                                                                \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess)                                                                       \
            return oe::fail(OE_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                            __FILE__, __LINE__);                                                    \
    } while (0)

#define OE_TRY(expr)                      \
    do {                                  \
        oe_status _s = (expr);            \
        if (_s != OE_OK) return _s;       \
    } while (0)

using oe::AdeDev;
using oe::CpmlArgs;
using oe::DevBuf;
using oe::DftDev;
using oe::Grid;
using oe::MurDev;
using oe::Nf2ffCachePtr;
using oe::Nf2ffFace;
using oe::Tiles;

// ---------------------------------------------------------------- operator
struct MatOp {
    int kind;  // 0 material, 1 PEC, 2 lumped resistor
    double eps_r, mu_r, se, sm;
    size_t box[6];
    int dir;   // lumped
    double R;  // lumped
};

struct oe_operator {
    size_t nx, ny, nz;
    double d[3];
    double dt;
    std::vector<MatOp> ops;
};

// ------------------------------------------------------------ extensions
enum ExtKind { EXT_CPML = 1, EXT_MUR = 2, EXT_ADE = 3, EXT_PROBE = 4, EXT_DFT = 5, EXT_UPML = 6, EXT_HOST = 7 };

struct oe_extension {
    int kind = 0;
    oe_engine* eng = nullptr;
    // CPML
    CpmlArgs cpml{};
    DevBuf cpml_mem;    // psi arrays
    DevBuf cpml_tab;
    size_t psi_count[6] = {0, 0, 0, 0, 0, 0};
    // second buffers of the z-face H psi (faces 4, 5) for fused steps, which
    // read the old psi and write the new one; swapped with cpml.ph_* per step
    float* zalt_pos[2] = {nullptr, nullptr};
    float* zalt_neg[2] = {nullptr, nullptr};
    // Mur
    std::vector<MurDev> mur_faces;
    DevBuf mur_mem;
    DevBuf mur_dev;     // device copy of mur_faces
    // UPML
    oe::UpmlDev upml{};
    DevBuf upml_mem;
    // ADE
    AdeDev ade{};
    DevBuf ade_mem;
    int ade_kmin = 0, ade_kmax = -1;  // k range of the cells (empty: kmax < kmin)
    std::vector<uint32_t> ade_cells;   // the cells' linear indices, sorted
    // probe
    uint32_t probe_idx = 0;
    DevBuf probe_mem;   // data
    DevBuf probe_cnt;   // uint64 counter
    uint64_t probe_cap = 0;
    // DFT (and lumped port)
    DftDev dft{};
    DevBuf dft_mem;
    uint64_t dft_count = 0;   // accumulated samples
    uint64_t dft_seen = 0;    // steps seen (time stride)
    uint32_t dft_stride = 1;
    std::vector<double> dft_freqs;
    bool is_port = false;
    // NF2FF box: faces whose phasors are consecutive DFT signals from nf_sig0[f]
    bool is_nf2ff = false;
    std::vector<Nf2ffFace> nf_faces;
    std::vector<size_t> nf_sig0;
    Nf2ffCachePtr nf_cache;
    // host-callback extension: hooks on fields copied to the host
    oe_host_callback_config host{};
    std::string host_reason;  // last reason check_termination gave
};

// ------------------------------------------------------------------ engine
struct KernelStat {
    uint64_t launches = 0;
    double ms = 0.0;
};

struct oe_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    Grid g{};
    double d[3] = {0, 0, 0};
    double dt = 0;
    float* F = nullptr;   // 6N, the current fields
    float* F2 = nullptr;  // 6N, the other buffer of fused steps (allocated on first use)
    int fused_mode = 0;   // 0: two half-steps; 1, 2: fused steps where the batch allows them
    float* Cf = nullptr;  // 12N
    // material table: row indices [E: N][H: N] and rows [E: cap x 6][H: cap x 6]
    uint16_t* mat = nullptr;
    float* mtab = nullptr;
    int n_mat[2] = {0, 0};  // rows of the E and H tables; 0: that half reads Cf
    uint8_t* mat8 = nullptr;  // [E: N][H: N] the indices as bytes, when a half has 2 to 256 rows
    uint64_t timestep = 0;
    Tiles tiles{};
    Tiles slab_tiles{};  // tiles of the CPML x/y slab launches (bx = 0: the interior's)
    int xcd_remap = 1;
    double peak_persistent = 0.0;
    oe_extension* cpml_owner = nullptr;
    std::vector<oe_extension*> exts;  // live extensions bound to this engine
    // energy
    double* partials = nullptr;
    double* energy_dev = nullptr;   // ring of [cap][2]
    size_t energy_dev_cap = 0;
    double* energy_host = nullptr;  // pinned
    // point ops
    DevBuf exc_samples;
    DevBuf step_dev;       // batch-relative step read by the kernels of a step
    int graph_steps = 16;  // steps per captured graph (0: no graphs)
    // CPML slab launches on a second stream beside the interior launch of
    // the same half-step (forked and joined with events; created on first
    // use): 1 on, 0 off.  Off by default: the slab work-groups only start as
    // interior ones retire and the cross-queue join adds about 15 us per
    // half-step (0.2815 against 0.2577 ms per step at 256^3).
    int overlap_slabs = 0;
    // z-face Mur applied by the E kernels where exact (oe_engine_run_batch
    // checks the conditions): 1 on, 0 off
    int fold_mur = 1;
    // dispersive extensions with disjoint cells in one pre and one post
    // launch per step: 1 on, 0 off
    int merge_ade = 1;
    // byte-wide row indices for tables of at most 256 rows: 1 on, 0 off
    int byte_index = 1;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DevBuf exc_table;
    DevBuf probe_table;
    // profiling: 0 off, n > 0 times every launch of every n-th step
    int profiling = 0;
    bool prof_step = false;
    std::map<std::string, KernelStat> stats;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // host copy of the six field arrays for host-callback hooks (pinned,
    // allocated on first use)
    float* host_fields = nullptr;
};

