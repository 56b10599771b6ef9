// oe_sweep.hip -- a parameter sweep: many structures (each an oe_engine with
// its CPML, lumped-port DFT and excitations) stepped together.
//
// The reference runs one Simulation at a time (Simulation::run,
// src/fdtd/simulation.rs:255-326, batching timesteps per call,
// src/fdtd/batch.rs:17-34); it has no multi-structure launcher.  Here every
// half-step of all active structures is one launch per kernel variant (the
// work-groups of all structures share the grid through a box table), the
// excitations, port DFTs and the step counter follow as one launch each, and
// the step sequence between two termination checks is captured once as a
// HIP graph and replayed.  Each structure stops on its own energy decay or at
// max_steps; stopped structures leave the tables.  Checks fall on the
// engines' absolute timesteps and the peak is kept in the engine, so a run
// can continue where an earlier one stopped with the same result as one run.

#include "oe_objects.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

using namespace oe;

namespace {

struct Item {
    oe_engine* eng;
    oe_extension* cpml;
    oe_extension* port;
    std::vector<SweepExc> exc;              // target filled at run time
    std::vector<std::vector<float>> samples;
    std::vector<size_t> exc_off;            // element offsets of the targets
    int vec = 4;
};

struct Group {      // one launch: all boxes of one (vec, axes) kind
    int vec, axes;
    std::vector<SweepBox> boxes;
    std::vector<int> wg_box;
    int n_wg = 0;
    const SweepBox* d_box = nullptr;  // device copies
    const int* d_wg = nullptr;
};

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

#define HIP_RET(expr)                          \
    do {                                       \
        const hipError_t e_ = (expr);          \
        if (e_ != hipSuccess) return e_;       \
    } while (0)

}  // namespace

struct oe_sweep {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Item> items;
    int ci = 8;            // planes per work-group along i
    int xcd_remap = 1;
    // largest material table of the run per half (0: some structure reads
    // its coefficient arrays, so every structure does)
    int rows_h = 0, rows_e = 0;
    double tune_ms[4] = {0, 0, 0, 0};
    int tune_ci[4] = {1, 2, 4, 8};
    // device state of a run
    DevBuf structs, tables, excs, exc_first, exc_serial, exc_samples, dfts, dft_of, ids, partials, energy, step;
    double* energy_host = nullptr;
    size_t energy_host_n = 0;
    std::vector<Group> groups;
    // per structure, the (timestep, E + H energy) of every check of the last run
    std::vector<std::vector<std::pair<uint64_t, double>>> history;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
    // The CPML slab launches of a half-step run beside the interior launch,
    // each on a side stream of its own (forked and joined by events; graph
    // branches under capture).  OE_SWEEP_OVERLAP=1: all slab launches on one
    // side stream; 0: every launch on the sweep stream.
    int overlap = 2;
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_steps = 0;
    }
    ~oe_sweep() {
        for (hipStream_t q : side)
            if (q) {
                (void)hipSetDevice(device);
                (void)hipStreamSynchronize(q);
            }
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        drop_graph();
        if (energy_host) (void)hipHostFree(energy_host);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (hipEvent_t q : ev_join)
            if (q) (void)hipEventDestroy(q);
        for (hipStream_t q : side)
            if (q) (void)hipStreamDestroy(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Boxes of one structure: with CPML the x-face slabs, the y-face slabs, the
// corner columns and the interior (z faces only), each with the kernel
// variant compiled for the faces it can hold; without CPML the whole grid
// (plain kernel).
void structure_boxes(const oe_sweep* sw, int sid, const Item& it, std::vector<Group>& groups) {
    const Grid& g = it.eng->g;
    const int vec = it.vec;
    const int nzv = ceil_div(g.nz, vec);
    SweepBox proto{};
    proto.sid = sid;
    proto.bx = std::min(64, nzv);
    proto.by = 256 / proto.bx;
    proto.n_kb = ceil_div(nzv, proto.bx);
    proto.ci = std::min(sw->ci, g.nx);
    auto add = [&](int axes, int i0, int i1, int j0, int j1) {
        if (i1 <= i0 || j1 <= j0) return;
        Group* G = nullptr;
        for (auto& q : groups)
            if (q.vec == vec && q.axes == axes) G = &q;
        if (!G) {
            groups.push_back(Group{vec, axes, {}, {}, 0, nullptr, nullptr});
            G = &groups.back();
        }
        SweepBox b = proto;
        b.i0 = i0; b.i1 = i1; b.j0 = j0; b.j1 = j1;
        // A box thinner in j than a work-group's rows (the y-face slabs and
        // corner columns: `layers` rows) stacks several i-chunks in one
        // work-group instead of leaving the surplus rows idle.
        const int rows = 256 / b.bx;
        b.by = std::min(rows, j1 - j0);
        b.ns = rows / b.by;
        b.n_jb = ceil_div(j1 - j0, b.by);
        const int n_ib = ceil_div(i1 - i0, b.ci * b.ns);
        b.first = G->n_wg;
        const int n = b.n_kb * b.n_jb * n_ib;
        for (int q = 0; q < n; ++q) G->wg_box.push_back((int)G->boxes.size());
        G->boxes.push_back(b);
        G->n_wg += n;
    };
    const CpmlArgs* cp = it.cpml ? &it.cpml->cpml : nullptr;
    if (!cp) {
        add(0, 0, g.nx, 0, g.ny);
        return;
    }
    const int m = cp->on_mask;
    const int ix0 = (m & 1) ? cp->L[0] : 0, ix1 = (m & 2) ? cp->g0[1] : g.nx;
    const int jy0 = (m & 4) ? cp->L[2] : 0, jy1 = (m & 8) ? cp->g0[3] : g.ny;
    // x slabs without the y-slab rows (x and z faces), y slabs without the
    // x-slab planes (y and z faces), the four corner columns (all faces):
    // the lighter kernels keep 3 waves per SIMD where the all-faces one has 2.
    const int all = AX_X | AX_Y | AX_Z;
    add(AX_X | AX_Z, 0, ix0, jy0, jy1);
    add(AX_X | AX_Z, ix1, g.nx, jy0, jy1);
    add(AX_Y | AX_Z, ix0, ix1, 0, jy0);
    add(AX_Y | AX_Z, ix0, ix1, jy1, g.ny);
    add(all, 0, ix0, 0, jy0);
    add(all, 0, ix0, jy1, g.ny);
    add(all, ix1, g.nx, 0, jy0);
    add(all, ix1, g.nx, jy1, g.ny);
    add((m & 48) ? AX_Z : 0, ix0, ix1, jy0, jy1);
}

// Launch tables of the active structures, uploaded into sw->tables.
oe_status build_tables(oe_sweep* sw, const std::vector<int>& active) {
    sw->groups.clear();
    for (int sid : active) structure_boxes(sw, sid, sw->items[sid], sw->groups);
    size_t nb = 0, nw = 0;
    std::vector<size_t> box_off, wg_off;
    for (auto& G : sw->groups) {
        box_off.push_back(nb);
        wg_off.push_back(nw);
        nb += G.boxes.size();
        nw += G.wg_box.size();
    }
    const size_t bbytes = (nb * sizeof(SweepBox) + 255) & ~(size_t)255;
    std::vector<char> host(bbytes + nw * sizeof(int));
    for (size_t q = 0; q < sw->groups.size(); ++q) {
        const Group& G = sw->groups[q];
        std::memcpy(host.data() + box_off[q] * sizeof(SweepBox), G.boxes.data(), G.boxes.size() * sizeof(SweepBox));
        std::memcpy(host.data() + bbytes + wg_off[q] * sizeof(int), G.wg_box.data(), G.wg_box.size() * sizeof(int));
    }
    if (sw->tables.bytes < host.size()) HIP_OK(sw->tables.alloc(host.size()));
    HIP_OK(hipMemcpyAsync(sw->tables.p, host.data(), host.size(), hipMemcpyHostToDevice, sw->stream));
    HIP_OK(hipStreamSynchronize(sw->stream));
    for (size_t q = 0; q < sw->groups.size(); ++q) {
        sw->groups[q].d_box = reinterpret_cast<const SweepBox*>(sw->tables.as<char>() + box_off[q] * sizeof(SweepBox));
        sw->groups[q].d_wg = reinterpret_cast<const int*>(sw->tables.as<char>() + bbytes + wg_off[q] * sizeof(int));
    }
    return OE_OK;
}

// step_inc (H half only): the first launch advances this step counter.  The
// launches of one half write disjoint cells and read only the other field,
// so with sw->overlap the slab groups (x/y CPML faces) go to side streams,
// forked from and joined back into the sweep stream around its interior
// launch.
hipError_t yee_half(oe_sweep* sw, bool h, uint64_t* step_inc = nullptr) {
    bool first = true;
    const bool ov = sw->overlap && sw->groups.size() > 1;
    const int n_side = sw->overlap == 2 ? 3 : 1;
    int used = 0;  // side streams forked so far
    if (ov) HIP_RET(hipEventRecord(sw->ev_fork, sw->stream));
    for (const auto& G : sw->groups) {
        SweepArgs a{};
        a.st = sw->structs.as<SweepStruct>();
        a.box = G.d_box;
        a.wg_box = G.d_wg;
        a.xcd_remap = sw->xcd_remap;
        a.step_inc = (h && first) ? step_inc : nullptr;
        first = false;
        hipStream_t st = sw->stream;
        if (ov && (G.axes & (AX_X | AX_Y)) != 0) {
            // slab group: the next side stream (or the one side stream)
            const int q = std::min(used, n_side - 1);
            if (q == used) {
                HIP_RET(hipStreamWaitEvent(sw->side[q], sw->ev_fork, 0));
                ++used;
            }
            st = sw->side[q];
        }
        hipError_t e = launch_sweep(h, a, G.n_wg, G.axes, G.vec, h ? sw->rows_h : sw->rows_e, st);
        if (e != hipSuccess) return e;
    }
    for (int q = 0; q < used; ++q) {
        HIP_RET(hipEventRecord(sw->ev_join[q], sw->side[q]));
        HIP_RET(hipStreamWaitEvent(sw->stream, sw->ev_join[q], 0));
    }
    return hipSuccess;
}

// Point and DFT tables of the active structures.
struct StepTables {
    int n_struct = 0;
    int n_dft = 0;
};

oe_status build_step_tables(oe_sweep* sw, const std::vector<int>& active, StepTables& t) {
    std::vector<SweepExc> ex;
    std::vector<int> first{0};
    std::vector<int> serial, dft_of;
    std::vector<SweepDft> dft;
    for (int sid : active) {
        const Item& it = sw->items[sid];
        for (const auto& e : it.exc) ex.push_back(e);
        first.push_back((int)ex.size());
        // sources on one cell apply in list order: then one thread does them all
        std::vector<const float*> cells;
        for (const auto& e : it.exc) cells.push_back(e.target);
        std::sort(cells.begin(), cells.end());
        serial.push_back(std::adjacent_find(cells.begin(), cells.end()) != cells.end() ? 1 : 0);
        dft_of.push_back(it.port ? (int)dft.size() : -1);
        if (it.port) {
            SweepDft d{};
            d.F = it.eng->F;
            d.N = it.eng->g.N;
            d.d = it.port->dft;
            d.sample0 = it.port->dft_count;
            d.t0 = (double)it.eng->timestep;
            dft.push_back(d);
        }
    }
    if (sw->excs.bytes < std::max<size_t>(ex.size(), 1) * sizeof(SweepExc))
        HIP_OK(sw->excs.alloc(std::max<size_t>(ex.size(), 1) * sizeof(SweepExc)));
    if (sw->exc_first.bytes < first.size() * sizeof(int)) HIP_OK(sw->exc_first.alloc(first.size() * sizeof(int)));
    const size_t nb_struct = std::max<size_t>(active.size(), 1) * sizeof(int);
    if (sw->exc_serial.bytes < nb_struct) HIP_OK(sw->exc_serial.alloc(nb_struct));
    if (sw->dft_of.bytes < nb_struct) HIP_OK(sw->dft_of.alloc(nb_struct));
    if (!active.empty()) {
        HIP_OK(hipMemcpyAsync(sw->exc_serial.p, serial.data(), serial.size() * sizeof(int), hipMemcpyHostToDevice,
                              sw->stream));
        HIP_OK(hipMemcpyAsync(sw->dft_of.p, dft_of.data(), dft_of.size() * sizeof(int), hipMemcpyHostToDevice,
                              sw->stream));
    }
    if (sw->dfts.bytes < std::max<size_t>(dft.size(), 1) * sizeof(SweepDft))
        HIP_OK(sw->dfts.alloc(std::max<size_t>(dft.size(), 1) * sizeof(SweepDft)));
    if (!ex.empty()) HIP_OK(hipMemcpyAsync(sw->excs.p, ex.data(), ex.size() * sizeof(SweepExc), hipMemcpyHostToDevice, sw->stream));
    HIP_OK(hipMemcpyAsync(sw->exc_first.p, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice, sw->stream));
    if (!dft.empty()) HIP_OK(hipMemcpyAsync(sw->dfts.p, dft.data(), dft.size() * sizeof(SweepDft), hipMemcpyHostToDevice, sw->stream));
    if (sw->ids.bytes < std::max<size_t>(active.size(), 1) * sizeof(int))
        HIP_OK(sw->ids.alloc(std::max<size_t>(active.size(), 1) * sizeof(int)));
    if (!active.empty())
        HIP_OK(hipMemcpyAsync(sw->ids.p, active.data(), active.size() * sizeof(int), hipMemcpyHostToDevice, sw->stream));
    HIP_OK(hipStreamSynchronize(sw->stream));
    t.n_struct = (int)active.size();
    t.n_dft = (int)dft.size();
    return OE_OK;
}

// One timestep of all active structures (no host synchronisation; capturable).
hipError_t sweep_step(oe_sweep* sw, const StepTables& t, double dt) {
    hipError_t e;
    // the first H launch advances the step counter (from -1: step st reads st)
    if ((e = yee_half(sw, true, sw->step.as<uint64_t>())) != hipSuccess) return e;
    if ((e = yee_half(sw, false)) != hipSuccess) return e;
    return launch_sweep_point(sw->excs.as<SweepExc>(), sw->exc_first.as<int>(), sw->exc_serial.as<int>(),
                              sw->dfts.as<SweepDft>(), sw->dft_of.as<int>(), t.n_struct, sw->step.as<uint64_t>(), dt,
                              sw->stream);
}

oe_status capture(oe_sweep* sw, const StepTables& t, double dt, int steps) {
    sw->drop_graph();
    HIP_OK(hipStreamBeginCapture(sw->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int q = 0; q < steps && err == hipSuccess; ++q) err = sweep_step(sw, t, dt);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(sw->stream, &g);
    if (err != hipSuccess || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(OE_ERR_DEVICE, "sweep graph capture failed: %s", hipGetErrorString(err != hipSuccess ? err : ec));
    }
    sw->graph = g;
    HIP_OK(hipGraphInstantiate(&sw->exec, sw->graph, nullptr, nullptr, 0));
    sw->graph_steps = steps;
    return OE_OK;
}

oe_status energies(oe_sweep* sw, int n) {
    if (n <= 0) return OE_OK;
    const size_t need = (size_t)n * 2 * energy_blocks() * sizeof(double);
    if (sw->partials.bytes < need) HIP_OK(sw->partials.alloc(need));
    if (sw->energy.bytes < (size_t)n * 2 * sizeof(double)) HIP_OK(sw->energy.alloc((size_t)n * 2 * sizeof(double)));
    if (sw->energy_host_n < (size_t)n) {
        if (sw->energy_host) (void)hipHostFree(sw->energy_host);
        sw->energy_host = nullptr;
        HIP_OK(hipHostMalloc(&sw->energy_host, (size_t)n * 2 * sizeof(double), hipHostMallocDefault));
        sw->energy_host_n = n;
    }
    HIP_OK(launch_sweep_energy(sw->structs.as<SweepStruct>(), sw->ids.as<int>(), n, sw->partials.as<double>(),
                               sw->energy.as<double>(), sw->stream));
    HIP_OK(hipMemcpyAsync(sw->energy_host, sw->energy.p, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost, sw->stream));
    HIP_OK(hipStreamSynchronize(sw->stream));
    return OE_OK;
}

}  // namespace

extern "C" {

oe_status oe_sweep_create(int device, oe_sweep** out) {
    if (!out) return fail(OE_ERR_CONFIG, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(OE_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(OE_ERR_CONFIG, "device %d out of range", device);
    std::unique_ptr<oe_sweep> sw(new oe_sweep());
    sw->device = device;
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&sw->stream, hipStreamNonBlocking));
    if (const char* v = getenv("OE_SWEEP_OVERLAP")) sw->overlap = std::min(std::max(atoi(v), 0), 2);
    if (sw->overlap) {
        HIP_OK(hipEventCreateWithFlags(&sw->ev_fork, hipEventDisableTiming));
        for (int q = 0; q < (sw->overlap == 2 ? 3 : 1); ++q) {
            HIP_OK(hipStreamCreateWithFlags(&sw->side[q], hipStreamNonBlocking));
            HIP_OK(hipEventCreateWithFlags(&sw->ev_join[q], hipEventDisableTiming));
        }
    }
    HIP_OK(sw->step.alloc(sizeof(uint64_t)));
    *out = sw.release();
    return OE_OK;
}

void oe_sweep_destroy(oe_sweep* sw) { delete sw; }

oe_status oe_sweep_add(oe_sweep* sw, oe_engine* e, oe_extension* cpml, oe_extension* port,
                       const oe_excitation* exc, size_t n_exc, size_t* index) {
    if (!sw || !e) return fail(OE_ERR_CONFIG, "null argument");
    if (n_exc && !exc) return fail(OE_ERR_CONFIG, "excitations is null");
    if (e->device != sw->device) return fail(OE_ERR_CONFIG, "engine on another device");
    for (const auto& it : sw->items)
        if (it.eng == e) return fail(OE_ERR_CONFIG, "engine already in the sweep");
    if (!sw->items.empty() && std::fabs(e->dt - sw->items[0].eng->dt) > 1e-12 * sw->items[0].eng->dt)
        return fail(OE_ERR_CONFIG, "all structures of a sweep share one timestep (dt differs)");
    if (cpml) {
        if (cpml->kind != EXT_CPML || cpml->eng != e) return fail(OE_ERR_CONFIG, "cpml is not this engine's CPML");
        if (!cpml->cpml.physical) return fail(OE_ERR_CONFIG, "a sweep runs the physical CPML mode only");
    }
    if (port) {
        if (port->kind != EXT_DFT || port->eng != e) return fail(OE_ERR_CONFIG, "port is not a DFT of this engine");
        if (port->dft.n_sig > 256) return fail(OE_ERR_CONFIG, "a sweep DFT holds at most 256 signals");
        if (port->dft_stride != 1) return fail(OE_ERR_CONFIG, "a sweep DFT samples every step");
    }
    Item it{};
    it.eng = e;
    it.cpml = cpml;
    it.port = port;
    it.vec = (e->g.nz % 4 == 0 && (!cpml || cpml->cpml.zvec)) ? 4 : 1;
    for (size_t q = 0; q < n_exc; ++q) {
        const oe_excitation& x = exc[q];
        if (x.i >= (uint32_t)e->g.nx || x.j >= (uint32_t)e->g.ny || x.k >= (uint32_t)e->g.nz || x.direction > 2)
            return fail(OE_ERR_CONFIG, "excitation %zu out of the grid", q);
        if (x.n_samples && !x.samples) return fail(OE_ERR_CONFIG, "excitation %zu has no samples", q);
        SweepExc s{};
        s.n_samples = x.n_samples;
        s.soft = x.soft ? 1u : 0u;
        it.exc.push_back(s);
        it.samples.emplace_back(x.samples, x.samples + x.n_samples);
        it.exc_off.push_back((size_t)x.direction * e->g.N + ((size_t)x.i * e->g.ny + x.j) * e->g.nz + x.k);
    }
    sw->items.push_back(std::move(it));
    if (index) *index = sw->items.size() - 1;
    return OE_OK;
}

oe_status oe_sweep_run(oe_sweep* sw, const oe_sweep_config* cfg, oe_sweep_entry* results) {
    if (!sw || !cfg || !results) return fail(OE_ERR_CONFIG, "null argument");
    if (sw->items.empty()) return fail(OE_ERR_CONFIG, "empty sweep");
    if (cfg->max_steps == 0) return fail(OE_ERR_CONFIG, "max_steps must be positive");
    HIP_OK(hipSetDevice(sw->device));
    const auto t0 = std::chrono::steady_clock::now();
    const int n = (int)sw->items.size();
    const uint64_t check = cfg->check_interval ? cfg->check_interval : 100;
    const double dt = sw->items[0].eng->dt;
    const uint64_t T0 = sw->items[0].eng->timestep;  // timestep of the run's first step
    for (auto& it : sw->items) {
        if (it.eng->timestep != T0)
            return fail(OE_ERR_CONFIG, "structures of a sweep run start at one timestep (%llu and %llu)",
                        (unsigned long long)T0, (unsigned long long)it.eng->timestep);
        HIP_OK(hipStreamSynchronize(it.eng->stream));
    }

    // static device data: structures, excitation samples
    std::vector<SweepStruct> st(n);
    size_t total = 0;
    for (int s = 0; s < n; ++s) {
        const Item& it = sw->items[s];
        st[s].g = it.eng->g;
        st[s].F = it.eng->F;
        st[s].Cf = it.eng->Cf;
        if (it.cpml) st[s].c = it.cpml->cpml;
        if (it.eng->mat) {
            st[s].mat_e = it.eng->mat;
            st[s].mat_h = it.eng->mat + it.eng->g.N;
            st[s].tab_e = it.eng->mtab;
            st[s].tab_h = it.eng->mtab + (size_t)MAX_MAT_ROWS * 6;
            st[s].n_mat_e = it.eng->n_mat[0];
            st[s].n_mat_h = it.eng->n_mat[1];
            if (it.eng->mat8) {
                if (st[s].n_mat_e >= 2 && st[s].n_mat_e <= 256) st[s].mat8_e = it.eng->mat8;
                if (st[s].n_mat_h >= 2 && st[s].n_mat_h <= 256) st[s].mat8_h = it.eng->mat8 + it.eng->g.N;
            }
        }
        for (const auto& v : it.samples) total += v.size();
    }
    sw->rows_e = sw->rows_h = 0;
    bool all_e = true, all_h = true;
    for (int s = 0; s < n; ++s) {
        all_e = all_e && st[s].n_mat_e > 0;
        all_h = all_h && st[s].n_mat_h > 0;
        sw->rows_e = std::max(sw->rows_e, st[s].n_mat_e);
        sw->rows_h = std::max(sw->rows_h, st[s].n_mat_h);
    }
    if (!all_e) sw->rows_e = 0;
    if (!all_h) sw->rows_h = 0;
    HIP_OK(sw->structs.alloc(n * sizeof(SweepStruct)));
    HIP_OK(hipMemcpyAsync(sw->structs.p, st.data(), n * sizeof(SweepStruct), hipMemcpyHostToDevice, sw->stream));
    HIP_OK(sw->exc_samples.alloc(std::max<size_t>(total, 1) * sizeof(float)));
    size_t off = 0;
    for (auto& it : sw->items)
        for (size_t q = 0; q < it.exc.size(); ++q) {
            it.exc[q].target = it.eng->F + it.exc_off[q];
            it.exc[q].samples = sw->exc_samples.as<float>() + off;
            if (!it.samples[q].empty())
                HIP_OK(hipMemcpyAsync(sw->exc_samples.as<float>() + off, it.samples[q].data(),
                                      it.samples[q].size() * sizeof(float), hipMemcpyHostToDevice, sw->stream));
            off += it.samples[q].size();
        }
    HIP_OK(hipMemsetAsync(sw->step.p, 0xff, sizeof(uint64_t), sw->stream));  // -1: the first step's H launch makes it 0

    std::vector<int> active(n);
    for (int s = 0; s < n; ++s) active[s] = s;
    std::vector<double> peak(n, 0.0);
    for (int s = 0; s < n; ++s) peak[s] = sw->items[s].eng->peak_persistent;
    sw->history.assign(n, {});
    for (int s = 0; s < n; ++s) {
        results[s] = oe_sweep_entry{};
        results[s].termination_reason = OE_TERM_STEPS_COMPLETED;
    }

    // Autotune the i-depth of the work-groups on the zero state: with zero
    // fields and no sources every update leaves the state exactly zero.
    if (cfg->autotune) {
        HIP_OK(sw->ids.alloc(n * sizeof(int)));
        HIP_OK(hipMemcpyAsync(sw->ids.p, active.data(), n * sizeof(int), hipMemcpyHostToDevice, sw->stream));
        OE_TRY(energies(sw, n));
        bool zero = true;
        for (int s = 0; s < n; ++s) zero = zero && sw->energy_host[2 * s] == 0.0 && sw->energy_host[2 * s + 1] == 0.0;
        if (zero) {
            hipEvent_t a, b;
            HIP_OK(hipEventCreate(&a));
            HIP_OK(hipEventCreate(&b));
            double best = 1e300;
            int best_ci = sw->ci;
            for (int q = 0; q < 4; ++q) {
                sw->ci = sw->tune_ci[q];
                OE_TRY(build_tables(sw, active));
                for (int w = 0; w < 2; ++w) {
                    HIP_OK(yee_half(sw, true));
                    HIP_OK(yee_half(sw, false));
                }
                HIP_OK(hipEventRecord(a, sw->stream));
                constexpr int TUNE_STEPS = 12;  // 6 left candidates within 1% to chance
                for (int w = 0; w < TUNE_STEPS; ++w) {
                    HIP_OK(yee_half(sw, true));
                    HIP_OK(yee_half(sw, false));
                }
                HIP_OK(hipEventRecord(b, sw->stream));
                HIP_OK(hipEventSynchronize(b));
                float ms = 0.f;
                HIP_OK(hipEventElapsedTime(&ms, a, b));
                sw->tune_ms[q] = ms / TUNE_STEPS;
                if (ms < best) {
                    best = ms;
                    best_ci = sw->tune_ci[q];
                }
            }
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            sw->ci = best_ci;
        }
    }

    OE_TRY(build_tables(sw, active));
    StepTables tab;
    OE_TRY(build_step_tables(sw, active, tab));
    uint64_t done = 0;
    bool dirty = false;  // tables changed since the last capture
    sw->drop_graph();
    while (!active.empty() && done < cfg->max_steps) {
        // Energy is sampled after the steps whose timestep is a multiple of
        // check (0, check, 2 check, ...), as a single engine with
        // sample_interval = check_interval does: from timestep 0, step 0
        // alone, then chunks of check steps.
        const uint64_t ts = T0 + done;                            // timestep of the next step
        const uint64_t next_check = (ts + check - 1) / check * check;  // first check at or after it
        const uint64_t chunk = std::min<uint64_t>(next_check - ts + 1, cfg->max_steps - done);
        if (cfg->use_graph && chunk == check && done > 0) {
            // the graph holds a divisor of the check interval (at most 25
            // steps) and is launched check / gsteps times: the active set
            // changes at many checks, and instantiating a short graph costs
            // a fraction of a long one
            int gsteps = 1;
            for (int q = 1; q <= 25; ++q)
                if (check % q == 0) gsteps = q;
            if (!sw->exec || dirty || sw->graph_steps != gsteps) {
                OE_TRY(capture(sw, tab, dt, gsteps));
                dirty = false;
            }
            for (uint64_t q = 0; q < chunk / gsteps; ++q) HIP_OK(hipGraphLaunch(sw->exec, sw->stream));
        } else {
            for (uint64_t q = 0; q < chunk; ++q) HIP_OK(sweep_step(sw, tab, dt));
        }
        done += chunk;
        const uint64_t cur = T0 + done - 1;  // timestep of the last step run
        if (cur % check != 0) continue;  // max_steps reached between checks
        OE_TRY(energies(sw, (int)active.size()));
        std::vector<int> keep;
        for (size_t q = 0; q < active.size(); ++q) {
            const int s = active[q];
            const double E = sw->energy_host[2 * q] + sw->energy_host[2 * q + 1];
            peak[s] = std::max(peak[s], E);
            sw->history[s].push_back({cur, E});
            results[s].final_energy = E;
            results[s].peak_energy = peak[s];
            bool stop = false;
            // src/fdtd/basic_engine.rs:128-141, with the engine's persistent peak
            if (cfg->has_energy_decay_db && cur > 0 && peak[s] > 0.0 && E > 0.0) {
                const double ddb = 10.0 * std::log10(E / peak[s]);
                if (ddb < -std::fabs(cfg->energy_decay_db)) {
                    stop = true;
                    results[s].termination_reason = OE_TERM_ENERGY_DECAY;
                    results[s].final_decay_db = ddb;
                }
            }
            if (stop) results[s].timesteps_executed = done;
            else keep.push_back(s);
        }
        if (keep.size() != active.size()) {
            active.swap(keep);
            if (!active.empty()) {
                OE_TRY(build_tables(sw, active));
                OE_TRY(build_step_tables(sw, active, tab));
                dirty = true;
            }
        }
    }
    HIP_OK(hipStreamSynchronize(sw->stream));
    for (int s : active) results[s].timesteps_executed = done;
    // Engines and port DFTs advance by the steps each structure ran.
    for (int s = 0; s < n; ++s) {
        Item& it = sw->items[s];
        it.eng->timestep += results[s].timesteps_executed;
        it.eng->peak_persistent = peak[s];
        if (it.port) {
            it.port->dft_count += results[s].timesteps_executed;
            it.port->dft_seen += results[s].timesteps_executed;
        }
    }
    sw->drop_graph();
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int s = 0; s < n; ++s) results[s].elapsed_seconds = wall;
    return OE_OK;
}

oe_status oe_sweep_tuning(const oe_sweep* sw, int* ci, double ms_per_step[4], int candidates[4]) {
    if (!sw) return fail(OE_ERR_CONFIG, "null argument");
    if (ci) *ci = sw->ci;
    for (int q = 0; q < 4; ++q) {
        if (ms_per_step) ms_per_step[q] = sw->tune_ms[q];
        if (candidates) candidates[q] = sw->tune_ci[q];
    }
    return OE_OK;
}

oe_status oe_sweep_energy_history(const oe_sweep* sw, size_t index, uint64_t* timesteps, double* energies,
                                  size_t capacity, size_t* n) {
    if (!sw) return fail(OE_ERR_CONFIG, "null argument");
    if (index >= sw->history.size()) return fail(OE_ERR_CONFIG, "structure %zu has no history (run the sweep first)", index);
    const auto& h = sw->history[index];
    if (n) *n = h.size();
    const size_t m = std::min(capacity, h.size());
    for (size_t q = 0; q < m; ++q) {
        if (timesteps) timesteps[q] = h[q].first;
        if (energies) energies[q] = h[q].second;
    }
    return OE_OK;
}

oe_status oe_sweep_set_ci(oe_sweep* sw, int ci) {
    if (!sw) return fail(OE_ERR_CONFIG, "null argument");
    if (ci < 1 || ci > 1024) return fail(OE_ERR_CONFIG, "ci out of range");
    sw->ci = ci;
    return OE_OK;
}

}  // extern "C"
