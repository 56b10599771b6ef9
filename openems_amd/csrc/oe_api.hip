// oe_api.hip -- host runtime behind include/oe/oe.h: operator, engine
// lifecycle, the per-step batch loop and the device-resident extensions.
//
// Mirrors Operator (src/fdtd/operator.rs), EngineImpl
// (src/fdtd/engine_impl.rs:24-71), the batch types (src/fdtd/batch.rs) and
// BasicEngine::run_batch (src/fdtd/basic_engine.rs:56-163).  No CPU compute
// fallback exists: every field update runs in a HIP kernel on the engine's
// stream; a missing or failing device is reported as OE_ERR_DEVICE.

#include "oe_objects.h"

#include <algorithm>
#include <chrono>
#include <complex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

using namespace oe;

// ------------------------------------------------------------------ errors
namespace {
thread_local std::string g_err;
}  // namespace

oe_status oe::fail(oe_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

oe_status set_device(int dev) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(OE_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (dev < 0 || dev >= n) return fail(OE_ERR_CONFIG, "device %d out of range (have %d)", dev, n);
    HIP_OK(hipSetDevice(dev));
    return OE_OK;
}

int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Automatic tile choice: 256-lane work-groups, as many lanes along k as the
// row holds (up to one wave of 64 float4 lanes), and i-chunks sized for
// about 16384 work-groups (64 per CU on 256 CUs), at least 4 planes deep;
// 2 planes on planes of more than 64 work-groups (384^2 and up) and on grids
// that 4 planes would cut into fewer than 1024 work-groups.  Measured with
// the material table, ms per step: 100^3 ci 1 / 2 / 4 = 0.0198 / 0.0197 /
// 0.0233; 256^3 ci 2 / 3 / 4 = 0.263 / 0.264 / 0.258; 384x384x256 with
// every extension ci 2 / 4 = 0.650 / 0.663; 512^3 ci 1 / 2 / 3 / 4 / 8 =
// 2.040 / 1.950 / 2.002 / 1.961 / 1.985.
Tiles auto_tiles(const Grid& g, int vec_req) {
    Tiles t{};
    t.vec = (vec_req == 1) ? 1 : ((g.nz % 4 == 0) ? 4 : 1);
    if (vec_req == 4 && g.nz % 4 != 0) t.vec = 1;
    const int nzv = (g.nz + t.vec - 1) / t.vec;
    t.bx = std::min(64, pow2ceil(std::max(nzv, 1)));
    t.by = 256 / t.bx;
    const int n_kb = (nzv + t.bx - 1) / t.bx;
    const int n_jb = (g.ny + t.by - 1) / t.by;
    const long plane_blocks = (long)n_kb * n_jb;
    long n_ib = (16384 + plane_blocks - 1) / plane_blocks;
    n_ib = std::max(1L, std::min(n_ib, (long)g.nx));
    int ci = (int)((g.nx + n_ib - 1) / n_ib);
    ci = std::max(ci, std::min(4, g.nx));
    const long wg_at_4 = plane_blocks * ((g.nx + 3) / 4);
    if (plane_blocks > 64 || wg_at_4 < 1024) ci = std::min(2, g.nx);
    t.ci = ci;
    return t;
}

struct Launch {
    YeeArgs a;
    int axes;
};

Box make_box(const Tiles& t, int i0, int i1, int j0, int j1) {
    Box b{};
    b.i0 = i0; b.i1 = i1; b.j0 = j0; b.j1 = j1;
    b.n_jb = (j1 - j0 + t.by - 1) / t.by;
    b.n_ib = (i1 - i0 + t.ci - 1) / t.ci;
    return b;
}

YeeArgs launch_args(const oe_engine* e, const Tiles& t, const std::vector<Box>& boxes) {
    YeeArgs a{};
    a.g = e->g;
    a.Fin = e->F;
    a.Fout = e->F;
    a.Cf = e->Cf;
    if (e->mat) {
        a.mat_e = e->mat;
        a.mat_h = e->mat + e->g.N;
        a.tab_e = e->mtab;
        a.tab_h = e->mtab + (size_t)MAX_MAT_ROWS * 6;
        a.n_mat_e = e->n_mat[0];
        a.n_mat_h = e->n_mat[1];
        if (e->mat8) {
            if (e->n_mat[0] >= 2 && e->n_mat[0] <= 256) a.mat8_e = e->mat8;
            if (e->n_mat[1] >= 2 && e->n_mat[1] <= 256) a.mat8_h = e->mat8 + e->g.N;
        }
    }
    a.t = t;
    const int nzv = (e->g.nz + t.vec - 1) / t.vec;
    a.n_kb = (nzv + t.bx - 1) / t.bx;
    a.nbox = (int)boxes.size();
    a.first[0] = 0;
    for (int b = 0; b < a.nbox; ++b) {
        a.box[b] = boxes[b];
        a.first[b + 1] = a.first[b] + a.n_kb * boxes[b].n_jb * boxes[b].n_ib;
    }
    a.xcd_remap = e->xcd_remap;
    return a;
}

// Launch plan of one half-step.  Without CPML: one launch of the plain kernel
// over the grid.  With CPML: the x and y face slabs (whole planes / rows) go
// to the all-faces kernel in one launch of up to four boxes, and the interior
// planes and rows -- whose only slab cells are the z faces at the ends of each
// row -- to the lighter z-faces kernel.  The boxes tile the grid exactly once.
std::vector<Launch> plan_launches(const oe_engine* e, const Tiles& t, const CpmlArgs* cp) {
    std::vector<Launch> out;
    const Grid& g = e->g;
    if (!cp) {
        out.push_back({launch_args(e, t, {make_box(t, 0, g.nx, 0, g.ny)}), 0});
        return out;
    }
    const int m = cp->on_mask;
    const int ix0 = (m & 1) ? cp->L[0] : 0, ix1 = (m & 2) ? cp->g0[1] : g.nx;
    const int jy0 = (m & 4) ? cp->L[2] : 0, jy1 = (m & 8) ? cp->g0[3] : g.ny;
    // One launch for all x/y slabs: split into x, y and corner launches
    // (lighter kernels, more waves) the slabs of a single 256^3 grid ran
    // slower (3 x 19 us against 31 us per half-step); the sweep, whose slab
    // launches carry every structure, does split them.
    std::vector<Box> slabs;
    Tiles ts = t;  // the interior's depth measured best for the slabs too (ci 1/2/4/8 at 256^3)
    if (e->slab_tiles.bx > 0) {
        ts = e->slab_tiles;
        if (ts.vec == 4 && g.nz % 4 != 0) ts.vec = 1;
    }
    if (ix0 > 0) slabs.push_back(make_box(ts, 0, ix0, 0, g.ny));
    if (ix1 < g.nx) slabs.push_back(make_box(ts, ix1, g.nx, 0, g.ny));
    if (ix1 > ix0) {
        if (jy0 > 0) slabs.push_back(make_box(ts, ix0, ix1, 0, jy0));
        if (jy1 < g.ny) slabs.push_back(make_box(ts, ix0, ix1, jy1, g.ny));
    }
    if (!slabs.empty()) out.push_back({launch_args(e, ts, slabs), AX_X | AX_Y | AX_Z});
    if (ix1 > ix0 && jy1 > jy0)
        out.push_back({launch_args(e, t, {make_box(t, ix0, ix1, jy0, jy1)}), (m & 48) ? AX_Z : 0});
    return out;
}

hipEvent_t take_event(oe_engine* e) {
    if (!e->event_pool.empty()) {
        hipEvent_t ev = e->event_pool.back();
        e->event_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

struct Timed {
    oe_engine* e;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(oe_engine* eng, const char* n) : e(eng), name(n) {
        if (e->profiling && e->prof_step) {
            a = take_event(e);
            b = take_event(e);
            if (a) (void)hipEventRecord(a, e->stream);
        }
    }
    ~Timed() {
        if (a && b) {
            (void)hipEventRecord(b, e->stream);
            e->pending.push_back({name, {a, b}});
        }
    }
};

void collect_timing(oe_engine* e) {
    for (auto& pr : e->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.second.first, pr.second.second) == hipSuccess) {
            auto& st = e->stats[pr.first];
            st.launches += 1;
            st.ms += ms;
        }
        e->event_pool.push_back(pr.second.first);
        e->event_pool.push_back(pr.second.second);
    }
    e->pending.clear();
}

oe_status check_region(const oe_operator* op, const size_t r[6]) {
    if (r[0] > r[1] || r[2] > r[3] || r[4] > r[5])
        return fail(OE_ERR_MATERIAL, "region start exceeds end");
    return OE_OK;
}

oe_status materialise(oe_engine* e, const oe_operator* op) {
    const Grid& g = e->g;
    const size_t N = g.N;
    const double dt = op->dt;
    // init_e/h_coefficients (src/fdtd/operator.rs:92-141)
    for (int d = 0; d < 3; ++d) {
        HIP_OK(launch_fill(e->Cf + (size_t)d * N, N, 1.0f, e->stream));
        HIP_OK(launch_fill(e->Cf + (size_t)(3 + d) * N, N, (float)(dt / (EPS0 * op->d[d])), e->stream));
        HIP_OK(launch_fill(e->Cf + (size_t)(6 + d) * N, N, 1.0f, e->stream));
        HIP_OK(launch_fill(e->Cf + (size_t)(9 + d) * N, N, (float)(-(dt / (MU0 * op->d[d]))), e->stream));
    }
    // The region writes in list order (later ones win where boxes meet),
    // cut into batches of pairwise-disjoint boxes: a box meeting one of the
    // current batch starts the next, so each batch is one launch and the
    // order of writes to any cell is kept.
    std::vector<BoxFill> fills;
    auto put = [&](const int bx[6], unsigned mask, const float v[12]) {
        if (bx[1] <= bx[0] || bx[3] <= bx[2] || bx[5] <= bx[4]) return;
        BoxFill f{};
        for (int q = 0; q < 6; ++q) f.box[q] = bx[q];
        f.mask = mask;
        for (int q = 0; q < 12; ++q) f.v[q] = v[q];
        fills.push_back(f);
    };
    for (size_t oi = 0; oi < op->ops.size(); ++oi) {
        const MatOp& m = op->ops[oi];
        int box[6];
        const size_t lim[6] = {0, (size_t)g.nx, 0, (size_t)g.ny, 0, (size_t)g.nz};
        for (int q = 0; q < 6; ++q) box[q] = (int)std::min(m.box[q], lim[q | 1]);
        float v[12] = {};
        if (m.kind == 2) {
            // Lumped resistor: per edge, the material in force before this op
            // plus the resistor's conductivity (python/openEMS/ports.py:177-181).
            const int d = m.dir, a = (d + 1) % 3, b = (d + 2) % 3;
            const size_t n[3] = {m.box[1] - m.box[0], m.box[3] - m.box[2], m.box[5] - m.box[4]};
            const double r_edge = m.R * (double)(n[a] * n[b]) / (double)n[d];
            const double sig_add = op->d[d] / (r_edge * op->d[a] * op->d[b]);
            for (int i = box[0]; i < box[1]; ++i)
                for (int j = box[2]; j < box[3]; ++j)
                    for (int k = box[4]; k < box[5]; ++k) {
                        double eps_r = 1.0, se = 0.0;
                        bool pec = false;
                        for (size_t pj = 0; pj < oi; ++pj) {
                            const MatOp& q = op->ops[pj];
                            if (q.kind == 2) continue;
                            if ((size_t)i >= q.box[0] && (size_t)i < q.box[1] && (size_t)j >= q.box[2] &&
                                (size_t)j < q.box[3] && (size_t)k >= q.box[4] && (size_t)k < q.box[5]) {
                                if (q.kind == 1) pec = true;
                                else { pec = false; eps_r = q.eps_r; se = q.se; }
                            }
                        }
                        if (pec) continue;
                        const double eps = eps_r * EPS0;
                        const double loss = (se + sig_add) * dt / (2.0 * eps);
                        const int cell[6] = {i, i + 1, j, j + 1, k, k + 1};
                        v[d] = (float)((1.0 - loss) / (1.0 + loss));
                        v[3 + d] = (float)((dt / eps) / (1.0 + loss) / op->d[d]);
                        put(cell, (1u << d) | (1u << (3 + d)), v);
                    }
            continue;
        }
        if (m.kind == 1) {  // src/fdtd/operator.rs:249-263
            put(box, 0x3fu, v);
            continue;
        }
        // src/fdtd/operator.rs:195-241 (computed in double, stored as float)
        const double eps = m.eps_r * EPS0;
        const double loss_e = m.se * dt / (2.0 * eps);
        const float ca = (float)((1.0 - loss_e) / (1.0 + loss_e));
        const double cbf = (dt / eps) / (1.0 + loss_e);
        const double mu = m.mu_r * MU0;
        const double loss_m = m.sm * dt / (2.0 * mu);
        const float da = (float)((1.0 - loss_m) / (1.0 + loss_m));
        const double dbf = -(dt / mu) / (1.0 + loss_m);
        for (int d = 0; d < 3; ++d) {
            v[d] = ca;
            v[3 + d] = (float)(cbf / op->d[d]);
            v[6 + d] = da;
            v[9 + d] = (float)(dbf / op->d[d]);
        }
        put(box, 0xfffu, v);
    }
    if (fills.empty()) return OE_OK;
    auto meets = [](const BoxFill& x, const BoxFill& y) {
        return x.box[0] < y.box[1] && y.box[0] < x.box[1] && x.box[2] < y.box[3] && y.box[2] < x.box[3] &&
               x.box[4] < y.box[5] && y.box[4] < x.box[5];
    };
    constexpr size_t MAX_BATCH = 1024;  // bounds the host's pairwise test
    std::vector<size_t> starts{0};
    for (size_t q = 1; q < fills.size(); ++q) {
        bool cut = q - starts.back() >= MAX_BATCH;
        for (size_t r = starts.back(); r < q && !cut; ++r) cut = meets(fills[r], fills[q]);
        if (cut) starts.push_back(q);
    }
    starts.push_back(fills.size());
    std::vector<int> wg_first(fills.size());
    std::vector<int> n_wg(starts.size() - 1);
    for (size_t b = 0; b + 1 < starts.size(); ++b) {
        int w = 0;
        for (size_t q = starts[b]; q < starts[b + 1]; ++q) {
            wg_first[q] = w;
            w += boxes_fill_blocks(fills[q]);
        }
        n_wg[b] = w;
    }
    DevBuf dev;
    const size_t fb = fills.size() * sizeof(BoxFill);
    HIP_OK(dev.alloc(fb + wg_first.size() * sizeof(int)));
    HIP_OK(hipMemcpy(dev.p, fills.data(), fb, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dev.as<char>() + fb, wg_first.data(), wg_first.size() * sizeof(int), hipMemcpyHostToDevice));
    const BoxFill* d_fills = dev.as<BoxFill>();
    const int* d_wg = reinterpret_cast<const int*>(dev.as<char>() + fb);
    for (size_t b = 0; b + 1 < starts.size(); ++b)
        HIP_OK(launch_boxes_fill(e->Cf, g, d_fills + starts[b], d_wg + starts[b], (int)(starts[b + 1] - starts[b]),
                                 n_wg[b], e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));  // the launches read `dev`
    return OE_OK;
}

oe_status engine_alloc(oe_engine* e, size_t nx, size_t ny, size_t nz, int device) {
    if (nx == 0 || ny == 0 || nz == 0) return fail(OE_ERR_GRID, "empty grid %zux%zux%zu", nx, ny, nz);
    if (nx > (1u << 30) || ny > (1u << 30) || nz > (1u << 30) || nx * ny * nz >= (1ull << 32))
        return fail(OE_ERR_GRID, "grid %zux%zux%zu exceeds 2^32 cells", nx, ny, nz);
    OE_TRY(set_device(device));
    e->device = device;
    e->g.nx = (int)nx; e->g.ny = (int)ny; e->g.nz = (int)nz;
    e->g.N = nx * ny * nz;
    HIP_OK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIP_OK(hipMalloc(&e->F, 6 * e->g.N * sizeof(float)));
    HIP_OK(hipMemsetAsync(e->F, 0, 6 * e->g.N * sizeof(float), e->stream));
    HIP_OK(hipMalloc(&e->Cf, 12 * e->g.N * sizeof(float)));
    HIP_OK(hipMalloc(&e->partials, 2 * energy_blocks() * sizeof(double)));
    e->energy_dev_cap = 4096;
    HIP_OK(hipMalloc(&e->energy_dev, 2 * e->energy_dev_cap * sizeof(double)));
    HIP_OK(hipHostMalloc(&e->energy_host, 2 * e->energy_dev_cap * sizeof(double), hipHostMallocDefault));
    e->tiles = auto_tiles(e->g, 0);
    if (const char* v = getenv("OE_XCD_REMAP")) e->xcd_remap = atoi(v) != 0;
    if (const char* v = getenv("OE_FUSED")) e->fused_mode = std::max(0, std::min(2, atoi(v)));
    if (const char* v = getenv("OE_GRAPH_STEPS")) e->graph_steps = std::max(0, atoi(v));
    if (const char* v = getenv("OE_OVERLAP_SLABS")) e->overlap_slabs = atoi(v) != 0;
    if (const char* v = getenv("OE_FOLD_MUR")) e->fold_mur = atoi(v) != 0;
    if (const char* v = getenv("OE_MERGE_ADE")) e->merge_ade = atoi(v) != 0;
    if (const char* v = getenv("OE_BYTE_INDEX")) e->byte_index = atoi(v) != 0;
    if (const char* v = getenv("OE_SLAB_TILES")) {  // "bx,by,ci,vec" of the CPML x/y slab launches
        Tiles t{};
        if (sscanf(v, "%d,%d,%d,%d", &t.bx, &t.by, &t.ci, &t.vec) == 4 && t.bx > 0 && t.bx <= 64 &&
            (t.bx & (t.bx - 1)) == 0 && t.bx * t.by == 256 && t.ci >= 1 && (t.vec == 1 || t.vec == 4))
            e->slab_tiles = t;
    }
    return OE_OK;
}

void drop_material_table(oe_engine* e) {
    if (e->mat) (void)hipFree(e->mat);
    if (e->mtab) (void)hipFree(e->mtab);
    if (e->mat8) (void)hipFree(e->mat8);
    e->mat = nullptr;
    e->mtab = nullptr;
    e->mat8 = nullptr;
    e->n_mat[0] = e->n_mat[1] = 0;
}

// Material table of both halves from the coefficient block (oe_coef.hip).
// A half with more than MAX_MAT_ROWS distinct coefficient sets keeps reading
// the arrays; with neither half tabulated the table is dropped.
oe_status build_material_tables(oe_engine* e) {
    drop_material_table(e);
    const size_t N = e->g.N;
    HIP_OK(hipMalloc(&e->mat, 2 * N * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&e->mtab, 2 * (size_t)MAX_MAT_ROWS * 6 * sizeof(float)));
    int n[2] = {0, 0};
    for (int h = 0; h < 2; ++h)
        HIP_OK(build_material_table(e->Cf, N, h == 1, MAX_MAT_ROWS, e->mat + (size_t)h * N,
                                    e->mtab + (size_t)h * MAX_MAT_ROWS * 6, &n[h], e->stream));
    if (n[0] == 0 && n[1] == 0) {
        drop_material_table(e);
        return OE_OK;
    }
    e->n_mat[0] = n[0];
    e->n_mat[1] = n[1];
    // halves of 2 to 256 rows also get their indices as bytes (1 B per cell
    // less per launch)
    const bool narrow[2] = {e->byte_index && n[0] >= 2 && n[0] <= 256, e->byte_index && n[1] >= 2 && n[1] <= 256};
    if (narrow[0] || narrow[1]) {
        HIP_OK(hipMalloc(&e->mat8, 2 * N));
        for (int h = 0; h < 2; ++h)
            if (narrow[h]) HIP_OK(narrow_material_index(e->mat + (size_t)h * N, e->mat8 + (size_t)h * N, N, e->stream));
    }
    return OE_OK;
}

bool material_table_default() {
    const char* v = getenv("OE_COEF_TABLE");
    return !(v && atoi(v) == 0);
}

void engine_free(oe_engine* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->side) (void)hipStreamSynchronize(e->side);
    // Extensions may outlive their engine: detach them so that later calls on
    // them report an error instead of touching the freed engine.
    for (oe_extension* x : e->exts) x->eng = nullptr;
    e->exts.clear();
    for (auto& pr : e->pending) {
        (void)hipEventDestroy(pr.second.first);
        (void)hipEventDestroy(pr.second.second);
    }
    for (auto ev : e->event_pool) (void)hipEventDestroy(ev);
    if (e->F) (void)hipFree(e->F);
    if (e->F2) (void)hipFree(e->F2);
    if (e->Cf) (void)hipFree(e->Cf);
    if (e->mat) (void)hipFree(e->mat);
    if (e->mtab) (void)hipFree(e->mtab);
    if (e->mat8) (void)hipFree(e->mat8);
    if (e->partials) (void)hipFree(e->partials);
    if (e->energy_dev) (void)hipFree(e->energy_dev);
    if (e->energy_host) (void)hipHostFree(e->energy_host);
    if (e->host_fields) (void)hipHostFree(e->host_fields);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
    if (e->side) (void)hipStreamDestroy(e->side);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* oe_last_error(void) { return g_err.c_str(); }
const char* oe_version(void) { return "openems_amd 0.1 (gfx950)"; }

oe_status oe_operator_create(size_t nx, size_t ny, size_t nz, const double delta[3], oe_operator** out) {
    if (!out || !delta) return fail(OE_ERR_CONFIG, "null argument");
    if (nx == 0 || ny == 0 || nz == 0) return fail(OE_ERR_GRID, "empty grid");
    for (int d = 0; d < 3; ++d)
        if (!(delta[d] > 0.0) || !std::isfinite(delta[d])) return fail(OE_ERR_GRID, "cell size must be positive");
    oe_operator* op = new oe_operator();
    op->nx = nx; op->ny = ny; op->nz = nz;
    for (int d = 0; d < 3; ++d) op->d[d] = delta[d];
    // Operator::calculate_timestep (src/fdtd/operator.rs:82-89)
    const double inv = 1.0 / (delta[0] * delta[0]) + 1.0 / (delta[1] * delta[1]) + 1.0 / (delta[2] * delta[2]);
    op->dt = 0.99 * (1.0 / (C0 * std::sqrt(inv)));
    *out = op;
    return OE_OK;
}

oe_status oe_operator_create_lines(const double* x, size_t nxl, const double* y, size_t nyl, const double* z,
                                   size_t nzl, oe_operator** out) {
    if (!x || !y || !z || nxl < 2 || nyl < 2 || nzl < 2) return fail(OE_ERR_GRID, "need at least 2 mesh lines per axis");
    const double* L[3] = {x, y, z};
    size_t n[3] = {nxl, nyl, nzl};
    double d[3];
    for (int a = 0; a < 3; ++a) {  // Grid::min_delta (src/geometry/mod.rs:113-118)
        double m = INFINITY;
        for (size_t i = 0; i + 1 < n[a]; ++i) m = std::min(m, std::fabs(L[a][i + 1] - L[a][i]));
        d[a] = m;
    }
    return oe_operator_create(nxl - 1, nyl - 1, nzl - 1, d, out);
}

void oe_operator_destroy(oe_operator* op) { delete op; }
double oe_operator_timestep(const oe_operator* op) { return op ? op->dt : 0.0; }
void oe_operator_dims(const oe_operator* op, size_t dims[3]) {
    if (!op || !dims) return;
    dims[0] = op->nx; dims[1] = op->ny; dims[2] = op->nz;
}

oe_status oe_operator_set_material(oe_operator* op, double eps_r, double mu_r, double sigma_e, double sigma_m,
                                   const size_t region[6]) {
    if (!op || !region) return fail(OE_ERR_CONFIG, "null argument");
    OE_TRY(check_region(op, region));
    if (!(eps_r > 0) || !(mu_r > 0)) return fail(OE_ERR_MATERIAL, "eps_r and mu_r must be positive");
    MatOp m{};
    m.kind = 0; m.eps_r = eps_r; m.mu_r = mu_r; m.se = sigma_e; m.sm = sigma_m;
    std::memcpy(m.box, region, sizeof(m.box));
    op->ops.push_back(m);
    return OE_OK;
}

oe_status oe_operator_set_pec(oe_operator* op, const size_t region[6]) {
    if (!op || !region) return fail(OE_ERR_CONFIG, "null argument");
    OE_TRY(check_region(op, region));
    MatOp m{};
    m.kind = 1;
    std::memcpy(m.box, region, sizeof(m.box));
    op->ops.push_back(m);
    return OE_OK;
}

oe_status oe_operator_add_lumped_element(oe_operator* op, int dir, const size_t region[6], double R) {
    if (!op || !region) return fail(OE_ERR_CONFIG, "null argument");
    OE_TRY(check_region(op, region));
    if (dir < 0 || dir > 2) return fail(OE_ERR_CONFIG, "direction must be 0, 1 or 2");
    if (!(R > 0.0) || !std::isfinite(R)) return fail(OE_ERR_MATERIAL, "lumped resistance must be positive");
    if (region[1] > op->nx || region[3] > op->ny || region[5] > op->nz)
        return fail(OE_ERR_GRID, "lumped element region exceeds the grid");
    for (int a = 0; a < 3; ++a)
        if (region[2 * a] == region[2 * a + 1]) return fail(OE_ERR_CONFIG, "lumped element region is empty");
    MatOp m{};
    m.kind = 2;
    m.dir = dir;
    m.R = R;
    std::memcpy(m.box, region, sizeof(m.box));
    op->ops.push_back(m);
    return OE_OK;
}

oe_status oe_engine_create(const oe_operator* op, int device, oe_engine** out) {
    if (!op || !out) return fail(OE_ERR_CONFIG, "null argument");
    oe_engine* e = new oe_engine();
    oe_status st = engine_alloc(e, op->nx, op->ny, op->nz, device);
    if (st == OE_OK) {
        for (int d = 0; d < 3; ++d) e->d[d] = op->d[d];
        e->dt = op->dt;
        st = materialise(e, op);
    }
    if (st == OE_OK && material_table_default()) st = build_material_tables(e);
    if (st == OE_OK && hipStreamSynchronize(e->stream) != hipSuccess) st = fail(OE_ERR_DEVICE, "engine setup failed");
    if (st != OE_OK) { engine_free(e); return st; }
    *out = e;
    return OE_OK;
}

oe_status oe_engine_create_from_arrays(size_t nx, size_t ny, size_t nz, const double delta[3], double dt,
                                       const float* const coeff[12], int device, oe_engine** out) {
    if (!coeff || !out || !delta) return fail(OE_ERR_CONFIG, "null argument");
    for (int w = 0; w < 12; ++w) if (!coeff[w]) return fail(OE_ERR_CONFIG, "coefficient array %d is null", w);
    oe_engine* e = new oe_engine();
    oe_status st = engine_alloc(e, nx, ny, nz, device);
    if (st == OE_OK) {
        for (int d = 0; d < 3; ++d) e->d[d] = delta[d];
        e->dt = dt;
        for (int w = 0; w < 12 && st == OE_OK; ++w)
            if (hipMemcpy(e->Cf + (size_t)w * e->g.N, coeff[w], e->g.N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                st = fail(OE_ERR_DEVICE, "coefficient upload failed");
    }
    if (st == OE_OK && material_table_default()) st = build_material_tables(e);
    if (st != OE_OK) { engine_free(e); return st; }
    *out = e;
    return OE_OK;
}

void oe_engine_destroy(oe_engine* eng) { engine_free(eng); }

uint64_t oe_engine_timestep(const oe_engine* eng) { return eng ? eng->timestep : 0; }

oe_status oe_engine_read_fields(oe_engine* e, float* const fields[6]) {
    if (!e || !fields) return fail(OE_ERR_CONFIG, "null argument");
    HIP_OK(hipSetDevice(e->device));
    for (int d = 0; d < 6; ++d)
        if (fields[d])
            HIP_OK(hipMemcpyAsync(fields[d], e->F + (size_t)d * e->g.N, e->g.N * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    return OE_OK;
}

oe_status oe_engine_write_fields(oe_engine* e, const float* const fields[6]) {
    if (!e || !fields) return fail(OE_ERR_CONFIG, "null argument");
    HIP_OK(hipSetDevice(e->device));
    for (int d = 0; d < 6; ++d)
        if (fields[d])
            HIP_OK(hipMemcpyAsync(e->F + (size_t)d * e->g.N, fields[d], e->g.N * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    return OE_OK;
}

oe_status oe_engine_graph_steps(oe_engine* e, int steps, int* current) {
    if (!e) return fail(OE_ERR_CONFIG, "null argument");
    if (steps < -1 || steps > 4096) return fail(OE_ERR_CONFIG, "graph steps must be -1 (query) or 0 .. 4096");
    if (steps >= 0) e->graph_steps = steps;
    if (current) *current = e->graph_steps;
    return OE_OK;
}

oe_status oe_engine_fused_steps(oe_engine* e, int mode, int* current) {
    if (!e) return fail(OE_ERR_CONFIG, "null argument");
    if (mode < -1 || mode > 2) return fail(OE_ERR_CONFIG, "mode must be -1, 0, 1 or 2");
    if (mode >= 0) e->fused_mode = mode;
    if (current) *current = e->fused_mode;
    return OE_OK;
}

oe_status oe_engine_material_table(oe_engine* e, int enable, int rows[2]) {
    if (!e) return fail(OE_ERR_CONFIG, "null argument");
    if (enable != -1 && enable != 0 && enable != 1) return fail(OE_ERR_CONFIG, "enable must be -1, 0 or 1");
    HIP_OK(hipSetDevice(e->device));
    if (enable == 0 && e->mat) {
        HIP_OK(hipStreamSynchronize(e->stream));
        drop_material_table(e);
    } else if (enable == 1 && !e->mat) {
        OE_TRY(build_material_tables(e));
        HIP_OK(hipStreamSynchronize(e->stream));
    }
    if (rows) {
        rows[0] = e->n_mat[0];
        rows[1] = e->n_mat[1];
    }
    return OE_OK;
}

oe_status oe_engine_read_coefficients(oe_engine* e, int which, float* out) {
    if (!e || !out) return fail(OE_ERR_CONFIG, "null argument");
    if (which < 0 || which >= 12) return fail(OE_ERR_CONFIG, "coefficient index %d out of range", which);
    HIP_OK(hipSetDevice(e->device));
    HIP_OK(hipMemcpyAsync(out, e->Cf + (size_t)which * e->g.N, e->g.N * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    return OE_OK;
}

oe_status oe_engine_reset(oe_engine* e) {
    if (!e) return fail(OE_ERR_CONFIG, "null argument");
    HIP_OK(hipSetDevice(e->device));
    HIP_OK(hipMemsetAsync(e->F, 0, 6 * e->g.N * sizeof(float), e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    e->timestep = 0;
    e->peak_persistent = 0.0;
    return OE_OK;
}

oe_status oe_engine_energy(oe_engine* e, double* ee, double* he) {
    if (!e) return fail(OE_ERR_CONFIG, "null argument");
    HIP_OK(hipSetDevice(e->device));
    HIP_OK(launch_energy(e->F, e->g.N, e->partials, energy_blocks(), e->energy_dev, e->stream));
    HIP_OK(hipMemcpyAsync(e->energy_host, e->energy_dev, 2 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    if (ee) *ee = e->energy_host[0];
    if (he) *he = e->energy_host[1];
    return OE_OK;
}

oe_status oe_engine_set_profiling(oe_engine* e, int on) {
    if (!e) return fail(OE_ERR_CONFIG, "null argument");
    e->profiling = on > 0 ? on : 0;
    if (e->profiling && e->event_pool.size() < 64) {
        // Create the event pool and time one empty interval now: the first
        // timed events of a process can cost tens of milliseconds of runtime
        // set-up, which must not land inside a measured batch.
        HIP_OK(hipSetDevice(e->device));
        while (e->event_pool.size() < 64) {
            hipEvent_t ev = nullptr;
            HIP_OK(hipEventCreate(&ev));
            e->event_pool.push_back(ev);
        }
        HIP_OK(hipEventRecord(e->event_pool[0], e->stream));
        HIP_OK(hipEventRecord(e->event_pool[1], e->stream));
        HIP_OK(hipEventSynchronize(e->event_pool[1]));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e->event_pool[0], e->event_pool[1]));
    }
    return OE_OK;
}

oe_status oe_engine_kernel_time(oe_engine* e, const char* name, uint64_t* launches, double* total_ms) {
    if (!e || !name) return fail(OE_ERR_CONFIG, "null argument");
    auto it = e->stats.find(name);
    if (launches) *launches = it == e->stats.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == e->stats.end() ? 0.0 : it->second.ms;
    return OE_OK;
}

oe_status oe_engine_set_tiles(oe_engine* e, int bx, int by, int ci, int vec) {
    if (!e) return fail(OE_ERR_CONFIG, "null argument");
    Tiles t = auto_tiles(e->g, vec);
    if (vec == 4 && e->g.nz % 4 != 0) return fail(OE_ERR_CONFIG, "vec=4 needs nz %% 4 == 0 (nz=%d)", e->g.nz);
    if (bx > 0) {
        if ((bx & (bx - 1)) != 0 || bx > 256) return fail(OE_ERR_CONFIG, "bx must be a power of two <= 256");
        t.bx = bx;
        t.by = 256 / bx;
    }
    if (by > 0) {
        if (by * t.bx != 256) return fail(OE_ERR_CONFIG, "bx * by must equal 256");
        t.by = by;
    }
    if (ci > 0) t.ci = std::min(ci, e->g.nx);
    e->tiles = t;
    return OE_OK;
}

oe_status oe_engine_get_tiles(const oe_engine* e, int tiles[4]) {
    if (!e || !tiles) return fail(OE_ERR_CONFIG, "null argument");
    tiles[0] = e->tiles.bx; tiles[1] = e->tiles.by; tiles[2] = e->tiles.ci; tiles[3] = e->tiles.vec;
    return OE_OK;
}

// ------------------------------------------------------------- waveforms
oe_status oe_waveform_evaluate(const oe_waveform* w, double t, double* value) {
    if (!w || !value) return fail(OE_ERR_CONFIG, "null argument");
    double v = 0.0;
    switch (w->type) {  // Excitation::evaluate (src/fdtd/excitation.rs:150-200)
        case OE_EXC_GAUSSIAN: {
            const double arg = (t - w->t0) / w->tau;
            v = std::exp(-arg * arg);
            break;
        }
        case OE_EXC_SINUSOIDAL: v = std::sin(2.0 * M_PI * w->frequency * t); break;
        case OE_EXC_GAUSSIAN_MODULATED: {
            const double arg = (t - w->t0) / w->tau;
            v = std::exp(-arg * arg) * std::sin(2.0 * M_PI * w->frequency * t);
            break;
        }
        case OE_EXC_DIRAC: v = std::fabs(t) < 1e-15 ? 1.0 : 0.0; break;
        case OE_EXC_STEP: v = t >= 0.0 ? 1.0 : 0.0; break;
        case OE_EXC_CUSTOM: {
            if (!w->times || !w->values || w->n_points == 0) return fail(OE_ERR_CONFIG, "custom waveform without points");
            const size_t n = w->n_points;
            if (t <= w->times[0]) v = w->values[0];
            else if (t >= w->times[n - 1]) v = w->values[n - 1];
            else {
                size_t i = 0;
                while (i < n - 1 && w->times[i + 1] < t) ++i;
                const double a = (t - w->times[i]) / (w->times[i + 1] - w->times[i]);
                v = w->values[i] + a * (w->values[i + 1] - w->values[i]);
            }
            break;
        }
        default: return fail(OE_ERR_CONFIG, "unknown waveform type %d", w->type);
    }
    *value = w->amplitude * v;
    return OE_OK;
}

oe_status oe_waveform_sample(const oe_waveform* w, uint64_t start, uint64_t n, double dt, float* out) {
    if (!w || (n && !out)) return fail(OE_ERR_CONFIG, "null argument");
    for (uint64_t s = 0; s < n; ++s) {
        double v;
        OE_TRY(oe_waveform_evaluate(w, (double)(start + s) * dt, &v));
        out[s] = (float)v;
    }
    return OE_OK;
}

// ------------------------------------------------------------ extensions
oe_status oe_ext_cpml_create(oe_engine* e, const oe_cpml_config* cfg, oe_extension** out) {
    if (!e || !cfg || !out) return fail(OE_ERR_CONFIG, "null argument");
    if (e->cpml_owner) return fail(OE_ERR_CONFIG, "engine already has a CPML extension");
    if (cfg->mode != OE_MODE_COMPAT && cfg->mode != OE_MODE_PHYSICAL) return fail(OE_ERR_CONFIG, "bad CPML mode");
    if (!(cfg->reflection_coeff > 0.0 && cfg->reflection_coeff < 1.0)) return fail(OE_ERR_CONFIG, "reflection_coeff must be in (0,1)");
    HIP_OK(hipSetDevice(e->device));
    std::unique_ptr<oe_extension> x(new oe_extension());
    x->kind = EXT_CPML;
    x->eng = e;
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    const double dmin = std::min(std::min(e->d[0], e->d[1]), e->d[2]);
    size_t lmax = 0;
    for (int f = 0; f < 6; ++f) lmax = std::max(lmax, cfg->layers[f]);
    // CpmlBoundaries::sigma_max (src/extensions/cpml.rs:68-72)
    const double dd = (double)(lmax ? lmax : 1) * dmin;
    const double m = cfg->grading_order;
    const double sigma_max = -(m + 1.0) * std::log(cfg->reflection_coeff) / (2.0 * Z0 * dd);
    CpmlArgs& c = x->cpml;
    c.physical = cfg->mode == OE_MODE_PHYSICAL;
    c.Lmax = (int)((std::max<size_t>(lmax, 1) + 3) & ~(size_t)3);  // 16-byte table rows
    for (int dir = 0; dir < 3; ++dir) {
        const size_t a0 = cfg->layers[2 * dir], a1 = cfg->layers[2 * dir + 1];
        const bool on0 = a0 > 0 && a0 < (size_t)n[dir], on1 = a1 > 0 && a1 < (size_t)n[dir];
        if (on0 && on1 && a0 + a1 > (size_t)n[dir])
            return fail(OE_ERR_CONFIG, "CPML faces of axis %d overlap (%zu + %zu layers > %d cells)", dir, a0, a1, n[dir]);
    }
    std::vector<float> tab((size_t)6 * CPML_TABLES * c.Lmax, 0.0f);
    size_t total = 0;
    size_t offs[6];
    for (int f = 0; f < 6; ++f) {
        const int dir = f / 2;
        const bool is_min = (f % 2) == 0;
        const size_t L = cfg->layers[f];
        offs[f] = total;
        if (!(L > 0 && L < (size_t)n[dir])) continue;  // Cpml::new (src/extensions/cpml.rs:262,283)
        c.on_mask |= 1 << f;
        c.L[f] = (int)L;
        c.g0[f] = is_min ? 0 : n[dir] - (int)L;
        const double delta = e->d[dir];
        const double width = (double)L * delta;
        for (size_t i = 0; i < L; ++i) {  // CpmlCoefficients::init (src/extensions/cpml.rs:98-177)
            const double de = is_min ? (double)(L - i) * delta : (double)(i + 1) * delta;
            const double dh = de - 0.5 * delta;
            const double re = std::min(std::max(de / width, 0.0), 1.0);
            const double rh = std::min(std::max(dh / width, 0.0), 1.0);
            const double se = sigma_max * std::pow(re, m), sh = sigma_max * std::pow(rh, m);
            const double ae_ = cfg->alpha_max * (1.0 - re), ah_ = cfg->alpha_max * (1.0 - rh);
            const double ke = 1.0 + (cfg->kappa_max - 1.0) * std::pow(re, m);
            const double kh = 1.0 + (cfg->kappa_max - 1.0) * std::pow(rh, m);
            const double be = std::exp(-(se / ke / EPS0 + ae_ / EPS0) * e->dt);
            const double bh = std::exp(-(sh / kh / EPS0 + ah_ / EPS0) * e->dt);
            const double sne = se / (EPS0 * ke), snh = sh / (EPS0 * kh);
            const double ane = ae_ / EPS0, anh = ah_ / EPS0;
            double a_e = std::fabs(sne + ane) > 1e-20 ? sne / (sne + ane) * (1.0 - be) : 0.0;
            double a_h = std::fabs(snh + anh) > 1e-20 ? snh / (snh + anh) * (1.0 - bh) : 0.0;
            if (c.physical) { a_e /= ke; a_h /= kh; }
            float* T = tab.data() + (size_t)f * CPML_TABLES * c.Lmax;
            T[TAB_BE * c.Lmax + i] = (float)be;
            T[TAB_BH * c.Lmax + i] = (float)bh;
            T[TAB_AE * c.Lmax + i] = (float)a_e;
            T[TAB_AH * c.Lmax + i] = (float)a_h;
            T[TAB_KE * c.Lmax + i] = (float)(1.0 / ke - 1.0);
            T[TAB_KH * c.Lmax + i] = (float)(1.0 / kh - 1.0);
        }
        size_t cnt = (size_t)L;
        for (int a = 0; a < 3; ++a) if (a != dir) cnt *= (size_t)n[a];
        cnt = (cnt + 63) & ~(size_t)63;  // 256-byte aligned sub-arrays
        x->psi_count[f] = cnt;
        total += 4 * cnt;
    }
    size_t zalt_off[2] = {0, 0};
    for (int f = 4; f < 6; ++f) {  // second z-face H psi buffers (fused steps)
        if (!(c.on_mask & (1 << f))) continue;
        zalt_off[f - 4] = total;
        total += 2 * x->psi_count[f];
    }
    const size_t dummy_off = total;  // 64 zeroed floats read by cells outside every slab
    total += 64;
    HIP_OK(x->cpml_mem.alloc(total * sizeof(float)));
    HIP_OK(hipMemsetAsync(x->cpml_mem.p, 0, x->cpml_mem.bytes, e->stream));
    c.dummy = x->cpml_mem.as<float>() + dummy_off;
    c.zvec = (e->g.nz % 4 == 0) ? 1 : 0;
    for (int f = 4; f < 6; ++f)
        if ((c.on_mask & (1 << f)) && c.L[f] % 4 != 0) c.zvec = 0;
    for (int f = 0; f < 6; ++f) {
        if (!(c.on_mask & (1 << f))) continue;
        float* base = x->cpml_mem.as<float>() + offs[f];
        const size_t cnt = x->psi_count[f];
        c.pe_pos[f] = base;
        c.pe_neg[f] = base + cnt;
        c.ph_pos[f] = base + 2 * cnt;
        c.ph_neg[f] = base + 3 * cnt;
        c.phw_pos[f] = c.ph_pos[f];
        c.phw_neg[f] = c.ph_neg[f];
        if (f >= 4) {
            x->zalt_pos[f - 4] = x->cpml_mem.as<float>() + zalt_off[f - 4];
            x->zalt_neg[f - 4] = x->cpml_mem.as<float>() + zalt_off[f - 4] + cnt;
        }
    }
    HIP_OK(x->cpml_tab.alloc(tab.size() * sizeof(float)));
    HIP_OK(hipMemcpyAsync(x->cpml_tab.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    c.tab = x->cpml_tab.as<float>();
    e->cpml_owner = x.get();
    e->exts.push_back(x.get());
    *out = x.release();
    return OE_OK;
}

oe_status oe_ext_mur_create(oe_engine* e, const int enabled[6], double phase_velocity, oe_extension** out) {
    if (!e || !enabled || !out) return fail(OE_ERR_CONFIG, "null argument");
    HIP_OK(hipSetDevice(e->device));
    std::unique_ptr<oe_extension> x(new oe_extension());
    x->kind = EXT_MUR;
    x->eng = e;
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    size_t total = 0;
    for (int f = 0; f < 6; ++f) {  // MurAbc::build_faces (src/extensions/mur_abc.rs:186-200)
        if (!enabled[f]) continue;
        MurDev m{};
        m.ny = f / 2;
        const bool is_max = (f % 2) == 1;
        m.nyp = (m.ny + 1) % 3;
        m.nypp = (m.ny + 2) % 3;
        if (n[m.ny] < 2) return fail(OE_ERR_GRID, "Mur face needs at least 2 cells along its normal");
        if (is_max) { m.line = n[m.ny] - 1; m.shift = n[m.ny] - 2; }
        else { m.line = 0; m.shift = 1; }
        m.s0 = n[m.nyp];
        m.s1 = n[m.nypp];
        const double c = phase_velocity > 0 ? phase_velocity : C0;
        const double cdt = c * e->dt;
        m.coeff = (float)((cdt - e->d[m.ny]) / (cdt + e->d[m.ny]));  // src/extensions/mur_abc.rs:138-142
        x->mur_faces.push_back(m);
        total += 2 * (((size_t)m.s0 * m.s1 + 63) & ~(size_t)63);
    }
    HIP_OK(x->mur_mem.alloc(std::max<size_t>(total, 64) * sizeof(float)));
    HIP_OK(hipMemsetAsync(x->mur_mem.p, 0, x->mur_mem.bytes, e->stream));
    size_t off = 0;
    for (auto& m : x->mur_faces) {
        const size_t cnt = ((size_t)m.s0 * m.s1 + 63) & ~(size_t)63;
        m.v_nyp = x->mur_mem.as<float>() + off;
        m.v_nypp = x->mur_mem.as<float>() + off + cnt;
        off += 2 * cnt;
    }
    HIP_OK(x->mur_dev.alloc(std::max<size_t>(x->mur_faces.size(), 1) * sizeof(MurDev)));
    if (!x->mur_faces.empty())
        HIP_OK(hipMemcpyAsync(x->mur_dev.p, x->mur_faces.data(), x->mur_faces.size() * sizeof(MurDev), hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    e->exts.push_back(x.get());
    *out = x.release();
    return OE_OK;
}

oe_status oe_ext_upml_create(oe_engine* e, const size_t layers[6], double m, double R, oe_extension** out) {
    if (!e || !layers || !out) return fail(OE_ERR_CONFIG, "null argument");
    if (!(R > 0.0 && R < 1.0)) return fail(OE_ERR_CONFIG, "reflection_coeff must be in (0, 1)");
    if (!(m >= 0.0)) return fail(OE_ERR_CONFIG, "grading_order must be non-negative");
    for (oe_extension* x : e->exts)
        if (x->kind == EXT_UPML) return fail(OE_ERR_CONFIG, "the engine already has a UPML");
    const size_t n[3] = {(size_t)e->g.nx, (size_t)e->g.ny, (size_t)e->g.nz};
    const double dt = e->dt;
    struct Reg { size_t st[3], sz[3]; int dir; bool is_min; };
    std::vector<Reg> regs;
    // Upml::build_regions (src/extensions/pml.rs:222-296)
    auto add = [&](size_t a0, size_t a1, size_t a2, size_t b0, size_t b1, size_t b2, int dir, bool is_min) {
        regs.push_back(Reg{{a0, a1, a2}, {b0, b1, b2}, dir, is_min});
    };
    const size_t nx = n[0], ny = n[1], nz = n[2];
    if (layers[0] > 0 && layers[0] < nx) add(0, 0, 0, layers[0] + 1, ny, nz, 0, true);
    if (layers[1] > 0 && layers[1] < nx) add(nx - layers[1] - 1, 0, 0, layers[1] + 1, ny, nz, 0, false);
    const size_t xs = layers[0] > 0 ? layers[0] + 1 : 0, xe = layers[1] > 0 ? nx - layers[1] - 1 : nx;
    if (layers[2] > 0 && layers[2] < ny && xs < xe) add(xs, 0, 0, xe - xs, layers[2] + 1, nz, 1, true);
    if (layers[3] > 0 && layers[3] < ny && xs < xe) add(xs, ny - layers[3] - 1, 0, xe - xs, layers[3] + 1, nz, 1, false);
    const size_t ys = layers[2] > 0 ? layers[2] + 1 : 0, ye = layers[3] > 0 ? ny - layers[3] - 1 : ny;
    if (layers[4] > 0 && layers[4] < nz && xs < xe && ys < ye) add(xs, ys, 0, xe - xs, ye - ys, layers[4] + 1, 2, true);
    if (layers[5] > 0 && layers[5] < nz && xs < xe && ys < ye)
        add(xs, ys, nz - layers[5] - 1, xe - xs, ye - ys, layers[5] + 1, 2, false);
    if (regs.empty()) return fail(OE_ERR_CONFIG, "no UPML region: every layer count is 0 or exceeds the grid");
    UpmlDev u{};
    u.n_reg = (int)regs.size();
    long total = 0;
    for (int r = 0; r < u.n_reg; ++r) {
        for (int d = 0; d < 3; ++d) { u.start[r][d] = (int)regs[r].st[d]; u.size[r][d] = (int)regs[r].sz[d]; }
        u.first[r] = total;
        total += (long)(regs[r].sz[0] * regs[r].sz[1] * regs[r].sz[2]);
    }
    u.first[u.n_reg] = total;
    // Upml::init_region_coefficients + calculate_kappa (src/extensions/pml.rs:299-461)
    std::vector<float> a((size_t)8 * 3 * total, 0.f);
    auto A = [&](int kind, int c, long q) -> float& { return a[((size_t)kind * 3 + c) * total + q]; };
    for (int r = 0; r < u.n_reg; ++r) {
        const Reg& g = regs[r];
        const int dir = g.dir;
        const size_t L = g.is_min ? layers[2 * dir] : layers[2 * dir + 1];
        const double delta = e->d[dir], width = (double)L * delta;
        const double smax = -(m + 1.0) * std::log(R) / (2.0 * Z0 * width);
        long q = u.first[r];
        for (size_t li = 0; li < g.sz[0]; ++li)
            for (size_t lj = 0; lj < g.sz[1]; ++lj)
                for (size_t lk = 0; lk < g.sz[2]; ++lk, ++q) {
                    const size_t pos[3] = {li + g.st[0], lj + g.st[1], lk + g.st[2]};
                    double kv[3], ki[3];
                    for (int c = 0; c < 3; ++c) {
                        double depth = 0.0;
                        if (g.is_min) {
                            if (pos[dir] <= L) depth = (double)(L - pos[dir]) * delta;
                        } else {
                            const size_t bs = n[dir] - L - 1;
                            if (pos[dir] >= bs) depth = (double)(pos[dir] - bs) * delta;
                        }
                        if (depth <= 0.0) { kv[c] = 0.0; ki[c] = 0.0; continue; }
                        kv[c] = smax * std::pow(depth / width, m);
                        const double di = depth + (c == dir ? -0.5 : 0.5) * delta;
                        double nd = di / width;
                        nd = nd < 0.0 ? 0.0 : (nd > 1.0 ? 1.0 : nd);
                        ki[c] = smax * std::pow(nd, m);
                    }
                    for (int c = 0; c < 3; ++c) {
                        const int npp = (c + 2) % 3;
                        if (kv[0] + kv[1] + kv[2] > 0.0) {
                            const double den = 2.0 * EPS0 + kv[npp] * dt;
                            A(0, c, q) = (float)((2.0 * EPS0 - kv[npp] * dt) / den);
                            A(1, c, q) = (float)((2.0 * EPS0 - kv[c] * dt) / den);
                            A(2, c, q) = (float)((2.0 * EPS0 + kv[c] * dt) / den);
                        } else {
                            A(0, c, q) = 1.f; A(1, c, q) = 0.f; A(2, c, q) = 1.f;
                        }
                        if (ki[0] + ki[1] + ki[2] > 0.0) {
                            const double den = 2.0 * EPS0 + ki[npp] * dt;
                            A(3, c, q) = (float)((2.0 * EPS0 - ki[npp] * dt) / den);
                            A(4, c, q) = (float)((2.0 * EPS0 - ki[c] * dt) / den);
                            A(5, c, q) = (float)((2.0 * EPS0 + ki[c] * dt) / den);
                        } else {
                            A(3, c, q) = 1.f; A(4, c, q) = 0.f; A(5, c, q) = 1.f;
                        }
                    }
                }
    }
    HIP_OK(hipSetDevice(e->device));
    std::unique_ptr<oe_extension> x(new oe_extension());
    x->kind = EXT_UPML;
    x->eng = e;
    HIP_OK(x->upml_mem.alloc(a.size() * sizeof(float)));
    HIP_OK(hipMemcpy(x->upml_mem.p, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice));
    u.a = x->upml_mem.as<float>();
    x->upml = u;
    e->exts.push_back(x.get());
    *out = x.release();
    return OE_OK;
}

oe_status oe_ext_upml_read(oe_extension* x, int region, int which, int component, float* out, size_t capacity,
                           size_t* n, size_t start[3], size_t size[3]) {
    if (!x || x->kind != EXT_UPML) return fail(OE_ERR_CONFIG, "not a UPML extension");
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    const UpmlDev& u = x->upml;
    if (region < 0 || region >= u.n_reg || which < 0 || which > 7 || component < 0 || component > 2)
        return fail(OE_ERR_CONFIG, "region, array or component out of range");
    const long total = u.first[u.n_reg];
    const size_t cnt = (size_t)(u.first[region + 1] - u.first[region]);
    if (n) *n = cnt;
    for (int d = 0; d < 3; ++d) {
        if (start) start[d] = (size_t)u.start[region][d];
        if (size) size[d] = (size_t)u.size[region][d];
    }
    if (out && capacity) {
        HIP_OK(hipSetDevice(x->eng->device));
        HIP_OK(hipStreamSynchronize(x->eng->stream));
        HIP_OK(hipMemcpy(out, u.a + ((size_t)which * 3 + component) * total + u.first[region],
                         std::min(capacity, cnt) * sizeof(float), hipMemcpyDeviceToHost));
    }
    return OE_OK;
}

oe_status oe_ext_ade_create(oe_engine* e, int kind, const double params[3], const uint32_t* positions, size_t n,
                            int mode, oe_extension** out) {
    if (!e || !params || !out || (n && !positions)) return fail(OE_ERR_CONFIG, "null argument");
    if (kind < 0 || kind > 2) return fail(OE_ERR_CONFIG, "unknown ADE kind %d", kind);
    if (mode != OE_MODE_COMPAT && mode != OE_MODE_PHYSICAL) return fail(OE_ERR_CONFIG, "bad ADE mode");
    if (n > (size_t)INT32_MAX) return fail(OE_ERR_CONFIG, "too many ADE cells");
    HIP_OK(hipSetDevice(e->device));
    std::unique_ptr<oe_extension> x(new oe_extension());
    x->kind = EXT_ADE;
    x->eng = e;
    std::vector<uint32_t> idx(n);
    for (size_t c = 0; c < n; ++c) {
        const uint32_t i = positions[3 * c], j = positions[3 * c + 1], k = positions[3 * c + 2];
        if (i >= (uint32_t)e->g.nx || j >= (uint32_t)e->g.ny || k >= (uint32_t)e->g.nz)
            return fail(OE_ERR_CONFIG, "ADE position %zu out of the grid", c);
        idx[c] = (uint32_t)(((size_t)i * e->g.ny + j) * e->g.nz + k);
        if (c == 0 || (int)k < x->ade_kmin) x->ade_kmin = (int)k;
        if ((int)k > x->ade_kmax) x->ade_kmax = (int)k;
    }
    x->ade_cells = idx;
    std::sort(x->ade_cells.begin(), x->ade_cells.end());
    AdeDev& a = x->ade;
    a.kind = kind; a.mode = mode; a.n = (int)n;
    const double dt = e->dt;
    if (kind == OE_ADE_LORENTZ) {  // src/extensions/dispersive.rs:134-159
        const double w0 = params[0], g = params[1], de = params[2];
        const double den = 1.0 + g * dt / 2.0;
        a.a = (float)((2.0 - w0 * w0 * dt * dt) / den);
        a.b = (float)((g * dt / 2.0 - 1.0) / den);
        a.c = (float)(EPS0 * de * w0 * w0 * dt * dt / den);
        a.pa = a.a; a.pb = a.b; a.pc = a.c;
    } else if (kind == OE_ADE_DRUDE) {  // :476-498
        const double wp = params[0], g = params[1];
        const double den = 1.0 + g * dt / 2.0;
        a.a = (float)((1.0 - g * dt / 2.0) / den);
        a.b = (float)(EPS0 * wp * wp * dt / den);
        a.c = (float)(dt / EPS0);
        a.pa = (float)((1.0 - g * dt / 2.0) / den);
        a.pb = (float)((EPS0 * wp * wp * dt / 2.0) / den);
    } else {  // Debye :570-592
        const double tau = params[0], de = params[1];
        const double den = 2.0 * tau + dt;
        a.a = (float)((2.0 * tau - dt) / den);
        a.b = (float)(2.0 * EPS0 * de * dt / den);
        a.c = (float)(dt / EPS0);
        a.pa = (float)((2.0 * tau - dt) / den);
        a.pb = (float)(EPS0 * de * dt / den);
    }
    a.dt_t = (float)dt;
    a.dtf = (float)dt;
    for (int d = 0; d < 3; ++d) a.delta[d] = e->d[d];
    const size_t cnt = (std::max<size_t>(n, 1) + 63) & ~(size_t)63;
    HIP_OK(x->ade_mem.alloc(cnt * sizeof(uint32_t) + 9 * cnt * sizeof(float)));
    HIP_OK(hipMemsetAsync(x->ade_mem.p, 0, x->ade_mem.bytes, e->stream));
    uint32_t* di = x->ade_mem.as<uint32_t>();
    float* fb = reinterpret_cast<float*>(di + cnt);
    a.idx = di;
    a.J = fb;
    a.Jold = fb + 3 * cnt;
    a.Eprev = fb + 6 * cnt;
    if (n) HIP_OK(hipMemcpyAsync(di, idx.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    e->exts.push_back(x.get());
    *out = x.release();
    return OE_OK;
}

oe_status oe_ext_field_probe_create(oe_engine* e, uint32_t i, uint32_t j, uint32_t k, oe_extension** out) {
    if (!e || !out) return fail(OE_ERR_CONFIG, "null argument");
    if (i >= (uint32_t)e->g.nx || j >= (uint32_t)e->g.ny || k >= (uint32_t)e->g.nz)
        return fail(OE_ERR_CONFIG, "probe position out of the grid");
    HIP_OK(hipSetDevice(e->device));
    std::unique_ptr<oe_extension> x(new oe_extension());
    x->kind = EXT_PROBE;
    x->eng = e;
    x->probe_idx = (uint32_t)(((size_t)i * e->g.ny + j) * e->g.nz + k);
    x->probe_cap = 0;
    HIP_OK(x->probe_cnt.alloc(sizeof(uint64_t)));
    HIP_OK(hipMemsetAsync(x->probe_cnt.p, 0, sizeof(uint64_t), e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    e->exts.push_back(x.get());
    *out = x.release();
    return OE_OK;
}

oe_status oe_ext_host_callback_create(oe_engine* e, const oe_host_callback_config* cfg, oe_extension** out) {
    if (!e || !cfg || !out) return fail(OE_ERR_CONFIG, "null argument");
    if (cfg->hook_mask & ~0x3Fu) return fail(OE_ERR_CONFIG, "hook_mask has bits above the six hook points");
    if (cfg->hook_mask && !cfg->hook) return fail(OE_ERR_CONFIG, "hook_mask selects hooks but hook is null");
    std::unique_ptr<oe_extension> x(new oe_extension());
    x->kind = EXT_HOST;
    x->eng = e;
    x->host = *cfg;
    e->exts.push_back(x.get());
    *out = x.release();
    return OE_OK;
}

oe_status oe_ext_host_callback_reason(oe_extension* x, char* out, size_t capacity, size_t* len) {
    if (!x) return fail(OE_ERR_CONFIG, "null argument");
    if (x->kind != EXT_HOST) return fail(OE_ERR_CONFIG, "not a host-callback extension");
    if (len) *len = x->host_reason.size();
    if (out && capacity) {
        const size_t n = std::min(capacity - 1, x->host_reason.size());
        std::memcpy(out, x->host_reason.data(), n);
        out[n] = 0;
    }
    return OE_OK;
}

oe_status oe_ext_dft_create(oe_engine* e, const oe_dft_config* cfg, oe_extension** out) {
    if (!e || !cfg || !out) return fail(OE_ERR_CONFIG, "null argument");
    const size_t n = cfg->n_signals, nf = cfg->n_freqs;
    if (n == 0 || n > (size_t)INT32_MAX || !cfg->offsets) return fail(OE_ERR_CONFIG, "need at least one signal");
    if (nf == 0 || nf > 4096 || !cfg->freqs) return fail(OE_ERR_CONFIG, "need 1..4096 frequencies");
    if (cfg->precision != OE_DFT_FP64 && cfg->precision != OE_DFT_FP32_KAHAN) return fail(OE_ERR_CONFIG, "bad DFT precision");
    const size_t nt = cfg->offsets[n];
    if (cfg->offsets[0] != 0 || (nt && (!cfg->component || !cfg->cell || !cfg->weight)))
        return fail(OE_ERR_CONFIG, "bad DFT term arrays");
    std::vector<uint8_t> is_h(n);
    for (size_t m = 0; m < n; ++m) {
        if (cfg->offsets[m + 1] < cfg->offsets[m] || cfg->offsets[m + 1] > nt) return fail(OE_ERR_CONFIG, "DFT offsets not monotone");
        int kind = -1;
        for (uint64_t t = cfg->offsets[m]; t < cfg->offsets[m + 1]; ++t) {
            if (cfg->component[t] > 5) return fail(OE_ERR_CONFIG, "DFT term component out of range");
            if (cfg->cell[t] >= e->g.N) return fail(OE_ERR_CONFIG, "DFT term cell out of the grid");
            const int h = cfg->component[t] >= 3;
            if (kind >= 0 && kind != h) return fail(OE_ERR_CONFIG, "signal %zu mixes E and H terms", m);
            kind = h;
        }
        is_h[m] = kind == 1;
    }
    HIP_OK(hipSetDevice(e->device));
    std::unique_ptr<oe_extension> x(new oe_extension());
    x->kind = EXT_DFT;
    x->eng = e;
    x->dft_freqs.assign(cfg->freqs, cfg->freqs + nf);
    x->dft_stride = cfg->time_stride > 1 ? cfg->time_stride : 1;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_off = al((n + 1) * 8), b_comp = al(std::max<size_t>(nt, 1) * 4), b_cell = al(std::max<size_t>(nt, 1) * 8);
    const size_t b_w = al(std::max<size_t>(nt, 1) * 8), b_h = al(n), b_f = al(nf * 8);
    const size_t b_acc = cfg->precision == OE_DFT_FP64 ? al(nf * n * 16) : al(nf * n * 16);
    const size_t b_ser = al(cfg->record_steps * n * 8);
    HIP_OK(x->dft_mem.alloc(b_off + b_comp + b_cell + b_w + b_h + b_f + b_acc + b_ser));
    char* base = x->dft_mem.as<char>();
    DftDev& d = x->dft;
    d.n_sig = (int)n;
    d.n_f = (int)nf;
    d.prec = cfg->precision;
    char* p = base;
    d.off = reinterpret_cast<const uint64_t*>(p); HIP_OK(hipMemcpy(p, cfg->offsets, (n + 1) * 8, hipMemcpyHostToDevice)); p += b_off;
    d.comp = reinterpret_cast<const uint32_t*>(p); if (nt) HIP_OK(hipMemcpy(p, cfg->component, nt * 4, hipMemcpyHostToDevice)); p += b_comp;
    d.cell = reinterpret_cast<const uint64_t*>(p); if (nt) HIP_OK(hipMemcpy(p, cfg->cell, nt * 8, hipMemcpyHostToDevice)); p += b_cell;
    d.w = reinterpret_cast<const double*>(p); if (nt) HIP_OK(hipMemcpy(p, cfg->weight, nt * 8, hipMemcpyHostToDevice)); p += b_w;
    d.is_h = reinterpret_cast<const uint8_t*>(p); HIP_OK(hipMemcpy(p, is_h.data(), n, hipMemcpyHostToDevice)); p += b_h;
    d.freqs = reinterpret_cast<const double*>(p); HIP_OK(hipMemcpy(p, cfg->freqs, nf * 8, hipMemcpyHostToDevice)); p += b_f;
    HIP_OK(hipMemset(p, 0, b_acc));
    if (cfg->precision == OE_DFT_FP64) d.acc64 = reinterpret_cast<double*>(p);
    else d.acc32 = reinterpret_cast<float*>(p);
    p += b_acc;
    d.series = cfg->record_steps ? reinterpret_cast<double*>(p) : nullptr;
    d.series_cap = cfg->record_steps;
    e->exts.push_back(x.get());
    *out = x.release();
    return OE_OK;
}

oe_status oe_ext_dft_read(oe_extension* x, double* re, double* im, uint64_t* n_acc) {
    if (!x || x->kind != EXT_DFT) return fail(OE_ERR_CONFIG, "not a DFT extension");
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    oe_engine* e = x->eng;
    HIP_OK(hipSetDevice(e->device));
    HIP_OK(hipStreamSynchronize(e->stream));
    const size_t cnt = (size_t)x->dft.n_f * x->dft.n_sig;
    const double scale = 2.0 * e->dt * x->dft_stride;  // pulse convention (python/openEMS/utilities.py:29-35)
    if (x->dft.prec == OE_DFT_FP64) {
        std::vector<double> h(2 * cnt);
        HIP_OK(hipMemcpy(h.data(), x->dft.acc64, h.size() * 8, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < cnt; ++q) {
            if (re) re[q] = scale * h[2 * q];
            if (im) im[q] = scale * h[2 * q + 1];
        }
    } else {
        std::vector<float> h(4 * cnt);
        HIP_OK(hipMemcpy(h.data(), x->dft.acc32, h.size() * 4, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < cnt; ++q) {
            if (re) re[q] = scale * (double)h[4 * q];
            if (im) im[q] = scale * (double)h[4 * q + 1];
        }
    }
    if (n_acc) *n_acc = x->dft_count;
    return OE_OK;
}

oe_status oe_ext_dft_read_series(oe_extension* x, double* out, size_t cap, size_t* n) {
    if (!x || x->kind != EXT_DFT) return fail(OE_ERR_CONFIG, "not a DFT extension");
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    oe_engine* e = x->eng;
    const size_t avail = (size_t)std::min<uint64_t>(x->dft_count, x->dft.series_cap);
    if (n) *n = avail;
    if (out && cap && x->dft.series) {
        HIP_OK(hipSetDevice(e->device));
        HIP_OK(hipStreamSynchronize(e->stream));
        HIP_OK(hipMemcpy(out, x->dft.series, std::min(cap, avail) * x->dft.n_sig * 8, hipMemcpyDeviceToHost));
    }
    return OE_OK;
}

oe_status oe_ext_lumped_port_create(oe_engine* e, int dir, const size_t r[6], int sign, const double* freqs,
                                    size_t n_freqs, int precision, uint64_t record_steps, oe_extension** out) {
    if (!e || !r || !out) return fail(OE_ERR_CONFIG, "null argument");
    if (dir < 0 || dir > 2 || (sign != 1 && sign != -1)) return fail(OE_ERR_CONFIG, "bad port direction or sign");
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    for (int q = 0; q < 3; ++q)
        if (r[2 * q] >= r[2 * q + 1] || r[2 * q + 1] > (size_t)n[q]) return fail(OE_ERR_GRID, "port region empty or outside the grid");
    const int a = (dir + 1) % 3, b = (dir + 2) % 3;
    const size_t lo[3] = {r[0], r[2], r[4]}, hi[3] = {r[1], r[3], r[5]};
    auto lin = [&](size_t p0, size_t p1, size_t p2) { return ((uint64_t)p0 * n[1] + p1) * n[2] + p2; };
    auto at = [&](size_t qd, size_t qa, size_t qb) {
        size_t pos[3];
        pos[dir] = qd; pos[a] = qa; pos[b] = qb;
        return lin(pos[0], pos[1], pos[2]);
    };
    std::vector<uint64_t> off{0}, cell;
    std::vector<uint32_t> comp;
    std::vector<double> w;
    // U: centre line of the cross-section, weight -sign * d_dir per edge
    const size_t ca = lo[a] + (hi[a] - lo[a] - 1) / 2, cb = lo[b] + (hi[b] - lo[b] - 1) / 2;
    for (size_t q = lo[dir]; q < hi[dir]; ++q) {
        comp.push_back((uint32_t)dir); cell.push_back(at(q, ca, cb)); w.push_back(-sign * e->d[dir]);
    }
    off.push_back(comp.size());
    // I: loop integral of H around every dir-edge of the cross-section at the middle index
    const size_t mid = lo[dir] + (hi[dir] - lo[dir]) / 2;
    for (size_t qa = lo[a]; qa < hi[a]; ++qa)
        for (size_t qb = lo[b]; qb < hi[b]; ++qb) {
            comp.push_back(3 + b); cell.push_back(at(mid, qa, qb)); w.push_back(sign * e->d[b]);
            if (qa >= 1) { comp.push_back(3 + b); cell.push_back(at(mid, qa - 1, qb)); w.push_back(-sign * e->d[b]); }
            comp.push_back(3 + a); cell.push_back(at(mid, qa, qb)); w.push_back(-sign * e->d[a]);
            if (qb >= 1) { comp.push_back(3 + a); cell.push_back(at(mid, qa, qb - 1)); w.push_back(sign * e->d[a]); }
        }
    off.push_back(comp.size());
    oe_dft_config c{};
    c.n_signals = 2;
    c.offsets = off.data();
    c.component = comp.data();
    c.cell = cell.data();
    c.weight = w.data();
    c.freqs = freqs;
    c.n_freqs = n_freqs;
    c.precision = precision;
    c.record_steps = record_steps;
    OE_TRY(oe_ext_dft_create(e, &c, out));
    (*out)->is_port = true;
    return OE_OK;
}

oe_status oe_ext_port_s11(oe_extension* x, double z_ref, double* s11_re, double* s11_im) {
    if (!x || x->kind != EXT_DFT || !x->is_port) return fail(OE_ERR_CONFIG, "not a lumped port extension");
    if (!(z_ref > 0.0)) return fail(OE_ERR_CONFIG, "z_ref must be positive");
    const size_t nf = x->dft_freqs.size();
    std::vector<double> re(2 * nf), im(2 * nf);
    OE_TRY(oe_ext_dft_read(x, re.data(), im.data(), nullptr));
    for (size_t f = 0; f < nf; ++f) {  // Port::CalcPort (python/openEMS/ports.py:136-139)
        const std::complex<double> U(re[2 * f], im[2 * f]), I(re[2 * f + 1], im[2 * f + 1]);
        const std::complex<double> inc = 0.5 * (U + I * z_ref);
        const std::complex<double> s = (U - inc) / inc;
        if (s11_re) s11_re[f] = s.real();
        if (s11_im) s11_im[f] = s.imag();
    }
    return OE_OK;
}

// ------------------------------------------------------------------ NF2FF
namespace {
void far_field_out(const Nf2ffOut& o, oe_far_field* out) {
    out->prad = o.prad;
    out->dmax = o.dmax;
    out->gemm_ms = o.gemm_ms;
    out->total_ms = o.total_ms;
    out->gemm_flop = o.gemm_flop;
    out->gemm_split_ms = o.gemm_split_ms;
    out->gemm_split_flop = o.gemm_split_flop;
}
Nf2ffOut far_field_in(const oe_far_field* out) {
    Nf2ffOut o{};
    o.et_re = out->e_theta_re;
    o.et_im = out->e_theta_im;
    o.ep_re = out->e_phi_re;
    o.ep_im = out->e_phi_im;
    o.et = out->e_theta;
    o.ep = out->e_phi;
    return o;
}
oe_status angles_ok(const double* theta, size_t nt, const double* phi, size_t np, double r, oe_far_field* out) {
    if (!theta || !phi || !out || nt == 0 || np == 0) return fail(OE_ERR_CONFIG, "need theta and phi angles");
    if (nt * np > (size_t)INT32_MAX / 16) return fail(OE_ERR_CONFIG, "too many angles");
    if (!(r > 0.0)) return fail(OE_ERR_CONFIG, "radius must be positive");
    return OE_OK;
}
}  // namespace

oe_status oe_ext_nf2ff_create(oe_engine* e, const size_t box[6], int directions, const double* freqs, size_t n_freqs,
                              int precision, uint32_t time_stride, oe_extension** out) {
    if (!e || !box || !out) return fail(OE_ERR_CONFIG, "null argument");
    const int n[3] = {e->g.nx, e->g.ny, e->g.nz};
    const size_t lo[3] = {box[0], box[2], box[4]}, hi[3] = {box[1], box[3], box[5]};
    for (int q = 0; q < 3; ++q)
        if (lo[q] < 1 || lo[q] >= hi[q] || hi[q] > (size_t)n[q] - 1)
            return fail(OE_ERR_GRID, "NF2FF box nodes must satisfy 1 <= lo < hi <= n-1 on every axis");
    if (directions == 0) directions = 0x3f;
    if (directions & ~0x3f) return fail(OE_ERR_CONFIG, "directions is a 6-bit face mask");
    auto lin = [&](const size_t p[3]) { return ((uint64_t)p[0] * n[1] + p[1]) * n[2] + p[2]; };
    std::vector<uint64_t> off{0}, cell;
    std::vector<uint32_t> comp;
    std::vector<double> w;
    std::vector<Nf2ffFace> faces;
    std::vector<size_t> sig0;
    size_t n_sig = 0;
    DftSurface surf{};
    surf.nx = n[0]; surf.ny = n[1]; surf.nz = n[2];
    for (int ax = 0; ax < 3; ++ax)
        for (int side = 0; side < 2; ++side) {
            if (!(directions & (1 << (2 * ax + side)))) continue;
            const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
            SurfFace& sf = surf.f[surf.n++];
            sf.ax = ax;
            sf.q = (int)(side ? hi[ax] : lo[ax]);
            sf.lo1 = (int)lo[a1]; sf.n1 = (int)(hi[a1] - lo[a1] + 1);
            sf.lo2 = (int)lo[a2]; sf.n2 = (int)(hi[a2] - lo[a2] + 1);
            sf.sig0 = (int)n_sig;
            Nf2ffFace f{};
            f.axis = ax;
            f.sign = side ? 1 : -1;
            const size_t qa = side ? hi[ax] : lo[ax];
            f.c0 = (double)qa * e->d[ax];
            for (size_t i = lo[a1]; i <= hi[a1]; ++i) f.c1.push_back((double)i * e->d[a1]);
            for (size_t i = lo[a2]; i <= hi[a2]; ++i) f.c2.push_back((double)i * e->d[a2]);
            sig0.push_back(n_sig);
            for (size_t i1 = lo[a1]; i1 <= hi[a1]; ++i1)
                for (size_t i2 = lo[a2]; i2 <= hi[a2]; ++i2) {
                    size_t p[3];
                    p[ax] = qa; p[a1] = i1; p[a2] = i2;
                    // E_t at the node: mean of the two edges that meet there
                    for (int c : {a1, a2}) {
                        for (int d = -1; d <= 0; ++d) {
                            size_t t[3] = {p[0], p[1], p[2]};
                            t[c] += d;
                            comp.push_back((uint32_t)c); cell.push_back(lin(t)); w.push_back(0.5);
                        }
                        off.push_back(comp.size());
                    }
                    // H_t at the node: mean of the four faces around it
                    for (int c : {a1, a2}) {
                        const int o = c == a1 ? a2 : a1;
                        for (int dn = -1; dn <= 0; ++dn)
                            for (int dt = -1; dt <= 0; ++dt) {
                                size_t t[3] = {p[0], p[1], p[2]};
                                t[ax] += dn;
                                t[o] += dt;
                                comp.push_back(3 + (uint32_t)c); cell.push_back(lin(t)); w.push_back(0.25);
                            }
                        off.push_back(comp.size());
                    }
                    n_sig += 4;
                }
            faces.push_back(std::move(f));
        }
    oe_dft_config c{};
    c.n_signals = n_sig;
    c.offsets = off.data();
    c.component = comp.data();
    c.cell = cell.data();
    c.weight = w.data();
    c.freqs = freqs;
    c.n_freqs = n_freqs;
    c.precision = precision;
    c.record_steps = 0;
    c.time_stride = time_stride;
    if (n_sig >= (size_t)INT32_MAX) return fail(OE_ERR_GRID, "NF2FF box of %zu signals exceeds 2^31", n_sig);
    OE_TRY(oe_ext_dft_create(e, &c, out));
    // the kernel forms the node values from the faces (the same terms, in
    // the same order, as the tables above); OE_SURFACE_DFT=0 reads the tables
    const char* sv = getenv("OE_SURFACE_DFT");
    if (!(sv && atoi(sv) == 0)) (*out)->dft.surf = surf;
    (*out)->is_nf2ff = true;
    (*out)->nf_faces = std::move(faces);
    (*out)->nf_sig0 = std::move(sig0);
    return OE_OK;
}

oe_status oe_ext_nf2ff_calculate(oe_extension* x, size_t fi, const double* theta, size_t nt, const double* phi,
                                 size_t np, double radius, const double center[3], oe_far_field* out) {
    if (!x || x->kind != EXT_DFT || !x->is_nf2ff) return fail(OE_ERR_CONFIG, "not an NF2FF extension");
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    OE_TRY(angles_ok(theta, nt, phi, np, radius, out));
    if (fi >= x->dft_freqs.size()) return fail(OE_ERR_CONFIG, "frequency index out of range");
    oe_engine* e = x->eng;
    HIP_OK(hipSetDevice(e->device));
    const double zero[3] = {0, 0, 0};
    const size_t base = fi * (size_t)x->dft.n_sig;
    for (size_t f = 0; f < x->nf_faces.size(); ++f) {
        Nf2ffFace& fc = x->nf_faces[f];
        const size_t q = base + x->nf_sig0[f];
        fc.src_kind = x->dft.prec == OE_DFT_FP64 ? 0 : 1;
        fc.src = fc.src_kind == 0 ? (const void*)(x->dft.acc64 + 2 * q) : (const void*)(x->dft.acc32 + 4 * q);
        fc.scale = 2.0 * e->dt * x->dft_stride;  // pulse convention (python/openEMS/utilities.py:29-35)
    }
    Nf2ffOut o = far_field_in(out);
    if (!x->nf_cache) x->nf_cache.reset(nf2ff_cache_new());
    HIP_OK(nf2ff_compute(x->nf_faces, x->dft_freqs[fi], theta, (int)nt, phi, (int)np, radius, center ? center : zero,
                         e->stream, x->nf_cache.get(), &o));
    far_field_out(o, out);
    return OE_OK;
}

struct oe_nf2ff {
    int device = 0;
    hipStream_t stream = nullptr;
    double freq = 0.0;
    std::vector<Nf2ffFace> faces;
    std::vector<std::unique_ptr<DevBuf>> bufs;
    Nf2ffCachePtr cache{nf2ff_cache_new()};
    ~oe_nf2ff() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

oe_status oe_nf2ff_create(int device, double frequency, oe_nf2ff** out) {
    if (!out) return fail(OE_ERR_CONFIG, "null argument");
    if (!(frequency > 0.0)) return fail(OE_ERR_CONFIG, "frequency must be positive");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(OE_ERR_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(OE_ERR_DEVICE, "device %d out of range", device);
    std::unique_ptr<oe_nf2ff> nf(new oe_nf2ff());
    nf->device = device;
    nf->freq = frequency;
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&nf->stream, hipStreamNonBlocking));
    *out = nf.release();
    return OE_OK;
}

void oe_nf2ff_destroy(oe_nf2ff* nf) { delete nf; }

oe_status oe_nf2ff_add_surface(oe_nf2ff* nf, int normal_dir, int sign, double position, const double* c1, size_t n1,
                               const double* c2, size_t n2, const double* e_t, const double* h_t) {
    if (!nf || !c1 || !c2 || !e_t || !h_t) return fail(OE_ERR_CONFIG, "null argument");
    if (normal_dir < 0 || normal_dir > 2 || (sign != 1 && sign != -1)) return fail(OE_ERR_CONFIG, "bad normal");
    if (n1 == 0 || n2 == 0 || n1 * n2 > (size_t)INT32_MAX / 16) return fail(OE_ERR_CONFIG, "bad surface size");
    const size_t nn = n1 * n2;
    std::vector<double> v(8 * nn);  // [n1][n2][4] complex: E_a1, E_a2, H_a1, H_a2
    for (size_t q = 0; q < nn; ++q)
        for (int c = 0; c < 2; ++c) {
            v[8 * q + 2 * c] = e_t[2 * (c * nn + q)];
            v[8 * q + 2 * c + 1] = e_t[2 * (c * nn + q) + 1];
            v[8 * q + 4 + 2 * c] = h_t[2 * (c * nn + q)];
            v[8 * q + 4 + 2 * c + 1] = h_t[2 * (c * nn + q) + 1];
        }
    HIP_OK(hipSetDevice(nf->device));
    std::unique_ptr<DevBuf> b(new DevBuf());
    HIP_OK(b->alloc(v.size() * sizeof(double)));
    HIP_OK(hipMemcpy(b->p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    Nf2ffFace f{};
    f.axis = normal_dir;
    f.sign = sign;
    f.c0 = position;
    f.c1.assign(c1, c1 + n1);
    f.c2.assign(c2, c2 + n2);
    f.src = b->p;
    f.src_kind = 0;
    f.scale = 1.0;
    nf->faces.push_back(std::move(f));
    nf->bufs.push_back(std::move(b));
    return OE_OK;
}

oe_status oe_nf2ff_calculate(oe_nf2ff* nf, const double* theta, size_t nt, const double* phi, size_t np,
                             double radius, const double center[3], oe_far_field* out) {
    if (!nf) return fail(OE_ERR_CONFIG, "null argument");
    OE_TRY(angles_ok(theta, nt, phi, np, radius, out));
    if (nf->faces.empty()) return fail(OE_ERR_CONFIG, "no surfaces added");
    const double zero[3] = {0, 0, 0};
    HIP_OK(hipSetDevice(nf->device));
    Nf2ffOut o = far_field_in(out);
    HIP_OK(nf2ff_compute(nf->faces, nf->freq, theta, (int)nt, phi, (int)np, radius, center ? center : zero,
                         nf->stream, nf->cache.get(), &o));
    far_field_out(o, out);
    return OE_OK;
}

static oe_status probe_reserve(oe_extension* x, uint64_t need) {
    if (need <= x->probe_cap) return OE_OK;
    oe_engine* e = x->eng;
    uint64_t cap = std::max<uint64_t>(need, std::max<uint64_t>(1024, 2 * x->probe_cap));
    DevBuf nb;
    HIP_OK(nb.alloc(cap * 6 * sizeof(double)));
    uint64_t cnt = 0;
    HIP_OK(hipMemcpyAsync(&cnt, x->probe_cnt.p, sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    if (cnt && x->probe_mem.p)
        HIP_OK(hipMemcpyAsync(nb.p, x->probe_mem.p, std::min(cnt, x->probe_cap) * 6 * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    std::swap(x->probe_mem.p, nb.p);
    std::swap(x->probe_mem.bytes, nb.bytes);
    x->probe_cap = cap;
    return OE_OK;
}

oe_status oe_ext_probe_read(oe_extension* x, double* out, size_t capacity, size_t* n_samples) {
    if (!x || x->kind != EXT_PROBE) return fail(OE_ERR_CONFIG, "not a probe extension");
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    oe_engine* e = x->eng;
    HIP_OK(hipSetDevice(e->device));
    uint64_t cnt = 0;
    HIP_OK(hipMemcpyAsync(&cnt, x->probe_cnt.p, sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    cnt = std::min(cnt, x->probe_cap);
    if (n_samples) *n_samples = cnt;
    if (out && capacity) {
        const size_t m = std::min<size_t>(capacity, cnt);
        if (m) HIP_OK(hipMemcpy(out, x->probe_mem.p, m * 6 * sizeof(double), hipMemcpyDeviceToHost));
    }
    return OE_OK;
}

oe_status oe_ext_cpml_read_psi(oe_extension* x, int face, int field, int which, float* out, size_t capacity, size_t* n) {
    if (!x || x->kind != EXT_CPML) return fail(OE_ERR_CONFIG, "not a CPML extension");
    if (face < 0 || face > 5 || field < 0 || field > 1 || which < 0 || which > 1) return fail(OE_ERR_CONFIG, "bad psi selector");
    const CpmlArgs& c = x->cpml;
    if (!(c.on_mask & (1 << face))) { if (n) *n = 0; return OE_OK; }
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    oe_engine* e = x->eng;
    size_t cnt = (size_t)c.L[face];
    const int dims[3] = {e->g.nx, e->g.ny, e->g.nz};
    for (int a = 0; a < 3; ++a) if (a != face / 2) cnt *= (size_t)dims[a];
    if (n) *n = cnt;
    const float* src = field == 0 ? (which == 0 ? c.pe_pos[face] : c.pe_neg[face]) : (which == 0 ? c.ph_pos[face] : c.ph_neg[face]);
    if (out && capacity) {
        HIP_OK(hipSetDevice(e->device));
        HIP_OK(hipMemcpyAsync(out, src, std::min(cnt, capacity) * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
    }
    return OE_OK;
}

oe_status oe_ext_ade_read_state(oe_extension* x, float* out, size_t capacity, size_t* n) {
    if (!x || x->kind != EXT_ADE) return fail(OE_ERR_CONFIG, "not an ADE extension");
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    const size_t cnt = (size_t)x->ade.n;
    if (n) *n = 3 * cnt;
    if (out && capacity) {
        oe_engine* e = x->eng;
        HIP_OK(hipSetDevice(e->device));
        for (int d = 0; d < 3; ++d) {
            const size_t m = std::min(cnt, capacity > d * cnt ? capacity - d * cnt : 0);
            if (m) HIP_OK(hipMemcpyAsync(out + d * cnt, x->ade.J + (size_t)d * cnt, m * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        }
        HIP_OK(hipStreamSynchronize(e->stream));
    }
    return OE_OK;
}

oe_status oe_extension_reset(oe_extension* x) {
    if (!x) return fail(OE_ERR_CONFIG, "null argument");
    if (!x->eng) return fail(OE_ERR_CONFIG, "the extension's engine was destroyed");
    oe_engine* e = x->eng;
    HIP_OK(hipSetDevice(e->device));
    if (x->kind == EXT_CPML) HIP_OK(hipMemsetAsync(x->cpml_mem.p, 0, x->cpml_mem.bytes, e->stream));
    if (x->kind == EXT_MUR) HIP_OK(hipMemsetAsync(x->mur_mem.p, 0, x->mur_mem.bytes, e->stream));
    if (x->kind == EXT_ADE) {
        const size_t cnt = (std::max<size_t>(x->ade.n, 1) + 63) & ~(size_t)63;
        HIP_OK(hipMemsetAsync(x->ade.J, 0, 9 * cnt * sizeof(float), e->stream));
    }
    if (x->kind == EXT_PROBE) HIP_OK(hipMemsetAsync(x->probe_cnt.p, 0, sizeof(uint64_t), e->stream));
    if (x->kind == EXT_UPML) {  // Upml::reset (src/extensions/pml.rs:660-667): the two flux arrays
        const long total = x->upml.first[x->upml.n_reg];
        HIP_OK(hipMemsetAsync(x->upml.a + (size_t)6 * 3 * total, 0, (size_t)2 * 3 * total * sizeof(float), e->stream));
    }
    if (x->kind == EXT_DFT) {
        void* acc = x->dft.prec == OE_DFT_FP64 ? (void*)x->dft.acc64 : (void*)x->dft.acc32;
        HIP_OK(hipMemsetAsync(acc, 0, (size_t)x->dft.n_f * x->dft.n_sig * 16, e->stream));
        x->dft_count = 0;
        x->dft_seen = 0;
    }
    HIP_OK(hipStreamSynchronize(e->stream));
    return OE_OK;
}

void oe_extension_destroy(oe_extension* x) {
    if (!x) return;
    if (x->eng) {
        oe_engine* e = x->eng;
        (void)hipSetDevice(e->device);
        (void)hipStreamSynchronize(e->stream);
        if (e->cpml_owner == x) e->cpml_owner = nullptr;
        e->exts.erase(std::remove(e->exts.begin(), e->exts.end(), x), e->exts.end());
    }
    delete x;
}

// ------------------------------------------------------------- run_batch
oe_status oe_engine_run_batch(oe_engine* e, const oe_batch_desc* b, oe_batch_result* r) {
    if (!e || !b || !r) return fail(OE_ERR_CONFIG, "null argument");
    if (b->n_excitations && !b->excitations) return fail(OE_ERR_CONFIG, "excitations is null");
    if (b->n_extensions && !b->extensions) return fail(OE_ERR_CONFIG, "extensions is null");
    HIP_OK(hipSetDevice(e->device));
    const auto t0 = std::chrono::steady_clock::now();
    std::memset(r, 0, sizeof(*r));

    const uint64_t num_steps = b->has_num_steps ? b->num_steps : (b->has_max_timesteps ? b->max_timesteps : UINT64_MAX);
    const uint64_t check_interval = b->check_interval ? b->check_interval : 100;

    // Sort extensions by hook.
    const CpmlArgs* cp = nullptr;
    oe_extension* cpx = nullptr;
    oe_extension* upml = nullptr;
    std::vector<oe_extension*> murs, ades, probes, dfts, hosts;
    for (size_t q = 0; q < b->n_extensions; ++q) {
        oe_extension* x = b->extensions[q];
        if (!x) return fail(OE_ERR_CONFIG, "extension %zu is null", q);
        if (!x->eng) return fail(OE_ERR_CONFIG, "extension %zu: its engine was destroyed", q);
        if (x->eng != e) return fail(OE_ERR_CONFIG, "extension %zu belongs to another engine", q);
        switch (x->kind) {
            case EXT_CPML:
                if (cp) return fail(OE_ERR_CONFIG, "more than one CPML extension in a batch");
                cp = &x->cpml;
                cpx = x;
                break;
            case EXT_UPML:
                if (upml) return fail(OE_ERR_CONFIG, "more than one UPML extension in a batch");
                upml = x;
                break;
            case EXT_MUR: murs.push_back(x); break;
            case EXT_ADE: ades.push_back(x); break;
            case EXT_PROBE: probes.push_back(x); break;
            case EXT_DFT: dfts.push_back(x); break;
            case EXT_HOST: hosts.push_back(x); break;
            default: return fail(OE_ERR_CONFIG, "unknown extension kind");
        }
    }

    // Excitation tables (ScheduledExcitation, src/fdtd/batch.rs:71-133).
    std::vector<PointExc> px(b->n_excitations);
    size_t total_samples = 0;
    for (size_t q = 0; q < b->n_excitations; ++q) {
        const oe_excitation& x = b->excitations[q];
        if (x.i >= (uint32_t)e->g.nx || x.j >= (uint32_t)e->g.ny || x.k >= (uint32_t)e->g.nz || x.direction > 2)
            return fail(OE_ERR_CONFIG, "excitation %zu out of the grid", q);
        if (x.n_samples && !x.samples) return fail(OE_ERR_CONFIG, "excitation %zu has no samples", q);
        total_samples += x.n_samples;
    }
    if (total_samples * sizeof(float) > e->exc_samples.bytes) HIP_OK(e->exc_samples.alloc(total_samples * sizeof(float) + 256));
    size_t off = 0;
    for (size_t q = 0; q < b->n_excitations; ++q) {
        const oe_excitation& x = b->excitations[q];
        px[q].idx = (uint32_t)(((size_t)x.i * e->g.ny + x.j) * e->g.nz + x.k);
        px[q].comp = x.direction;
        px[q].soft = x.soft ? 1u : 0u;
        px[q].samples = e->exc_samples.as<float>() + off;
        px[q].n_samples = x.n_samples;
        if (x.n_samples)
            HIP_OK(hipMemcpyAsync(e->exc_samples.as<float>() + off, x.samples, x.n_samples * sizeof(float), hipMemcpyHostToDevice, e->stream));
        off += x.n_samples;
    }
    if (px.size() * sizeof(PointExc) > e->exc_table.bytes) HIP_OK(e->exc_table.alloc(px.size() * sizeof(PointExc) + 64));
    if (!px.empty())
        HIP_OK(hipMemcpyAsync(e->exc_table.p, px.data(), px.size() * sizeof(PointExc), hipMemcpyHostToDevice, e->stream));

    // Probe tables.
    std::vector<PointProbe> pp(probes.size());
    const uint64_t reserve_steps = num_steps == UINT64_MAX ? 0 : num_steps;
    for (size_t q = 0; q < probes.size(); ++q) {
        oe_extension* x = probes[q];
        uint64_t cnt = 0;
        HIP_OK(hipMemcpyAsync(&cnt, x->probe_cnt.p, sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
        OE_TRY(probe_reserve(x, cnt + std::max<uint64_t>(reserve_steps, 1)));
        pp[q].idx = x->probe_idx;
        pp[q].out = x->probe_mem.as<double>();
        pp[q].capacity = x->probe_cap;
        pp[q].count = x->probe_cnt.as<uint64_t>();
    }
    if (pp.size() * sizeof(PointProbe) > e->probe_table.bytes) HIP_OK(e->probe_table.alloc(pp.size() * sizeof(PointProbe) + 64));
    if (!pp.empty())
        HIP_OK(hipMemcpyAsync(e->probe_table.p, pp.data(), pp.size() * sizeof(PointProbe), hipMemcpyHostToDevice, e->stream));
    Tiles tl = e->tiles;
    if (cp && !cp->zvec && tl.vec == 4) tl = auto_tiles(e->g, 1);  // z slabs not 4-aligned: scalar lanes
    const std::vector<Launch> plan = plan_launches(e, tl, cp);
    // Fused steps (one pass per step over the plan's interior box) when no
    // extension hooks in between the half-steps and the grid suits the
    // fused kernel; the x/y CPML slabs keep their two launches around it.
    // Host-callback extensions run between the launches of a step, so their
    // batches launch every step directly, in the plain two-half-step order.
    const bool host = !hosts.empty();
    const bool fused = e->fused_mode && !host && !upml && murs.empty() && ades.empty() && tl.vec == 4 &&
                       fused_supported(e->g) && (e->fused_mode == 2 || e->g.nz >= 128) && (!cp || cp->zvec);
    // Sources are added by the E kernels themselves when nothing acts between
    // the E update and them (no ADE, Mur or UPML post-E hooks), when every
    // source cell is one the E update writes (i, j >= 1), the list is short
    // and the steps are not fused (the fused kernel leaves sources to the
    // point kernel); the point kernel then only records probes.
    bool fold_sources = !fused && !host && !px.empty() && px.size() <= (size_t)MAX_FOLD_SOURCES && murs.empty() &&
                        ades.empty() && !upml;
    for (size_t q = 0; q < b->n_excitations && fold_sources; ++q)
        if (b->excitations[q].i == 0 || b->excitations[q].j == 0) fold_sources = false;
    if (const char* v = getenv("OE_FOLD_SOURCES")) fold_sources = fold_sources && atoi(v) != 0;
    const bool point_ops = (!px.empty() && !fold_sources) || !pp.empty();
    FusedArgs fa{};
    std::vector<Launch> slabs;
    if (fused) {
        if (!e->F2) HIP_OK(hipMalloc(&e->F2, 6 * e->g.N * sizeof(float)));
        const Launch* inner = nullptr;
        for (const Launch& L : plan) {
            if (L.axes & (AX_X | AX_Y)) slabs.push_back(L);
            else inner = &L;
        }
        fa.g = e->g;
        fa.Cf = e->Cf;
        if (e->mat && e->n_mat[0] > 0 && e->n_mat[1] > 0) {
            fa.mat_e = e->mat;
            fa.mat_h = e->mat + e->g.N;
            fa.tab_e = e->mtab;
            fa.tab_h = e->mtab + (size_t)MAX_MAT_ROWS * 6;
            fa.n_mat_e = e->n_mat[0];
            fa.n_mat_h = e->n_mat[1];
        }
        if (inner) {
            const Box& B = inner->a.box[0];
            fa.i0 = B.i0; fa.i1 = B.i1; fa.j0 = B.j0; fa.j1 = B.j1;
        }
        fa.n_jb = inner ? (fa.j1 - fa.j0 + FUSED_ROWS - 1) / FUSED_ROWS : 0;
        // about 2048 work-groups, 4 to 32 planes each
        const int ni = fa.i1 - fa.i0;
        int ci = fa.n_jb ? (int)(((long)ni * fa.n_jb + 2047) / 2048) : 1;
        if (const char* v = getenv("OE_FUSED_CI")) ci = atoi(v);
        fa.ci = std::max(1, std::min(std::max(ci, 4), 32));
        fa.n_ib = inner ? (ni + fa.ci - 1) / fa.ci : 0;
        fa.xcd_remap = e->xcd_remap;
        fa.kind = 2;
        if (const char* v = getenv("OE_FUSED_KIND")) fa.kind = atoi(v) == 1 ? 1 : 2;
        if (fa.kind == 2 && e->mat8 && e->n_mat[0] >= 2 && e->n_mat[0] <= 256) fa.mat8_e = e->mat8;
    }
    double peak = b->persistent_peak ? e->peak_persistent : 0.0;
    std::vector<uint64_t> sample_steps;  // timesteps whose energy is pending in energy_dev
    size_t n_samples_total = 0;
    double last_total = 0.0;
    const bool need_decay = b->has_energy_decay_db;

    auto drain = [&](void) -> oe_status {
        if (sample_steps.empty()) return OE_OK;
        HIP_OK(hipMemcpyAsync(e->energy_host, e->energy_dev, 2 * sample_steps.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
        for (size_t q = 0; q < sample_steps.size(); ++q) {
            const double ee = e->energy_host[2 * q], he = e->energy_host[2 * q + 1];
            if (b->track_peak) peak = std::max(peak, ee + he);
            last_total = ee + he;
            if (b->energy_samples && n_samples_total < b->energy_samples_capacity) {
                oe_energy_sample& s = b->energy_samples[n_samples_total];
                s.timestep = sample_steps[q];
                s.e_energy = ee;
                s.h_energy = he;
                s.total_energy = ee + he;
            }
            ++n_samples_total;
        }
        sample_steps.clear();
        return OE_OK;
    };

    uint64_t executed = 0;
    int reason = OE_TERM_STEPS_COMPLETED;
    double final_db = 0.0;

    // The step index seen by the kernels lives on the device (batch-relative),
    // so a run of steps without host work in between (no energy sample, decay
    // check or timed step) is replayed from one captured graph.  The DFTs
    // then form their sample times from the batch's first step.
    if (e->step_dev.bytes < sizeof(uint64_t)) HIP_OK(e->step_dev.alloc(sizeof(uint64_t)));
    uint64_t* step_dev = e->step_dev.as<uint64_t>();  // reset below, once the step scheme is known
    const uint64_t t_batch0 = e->timestep;
    bool dev_dft = true;
    std::vector<uint64_t> dft0(dfts.size());
    for (size_t q = 0; q < dfts.size(); ++q) {
        dev_dft = dev_dft && dfts[q]->dft_stride == 1;
        dft0[q] = dfts[q]->dft_count;
    }

    // First-order Mur on the z faces goes into the E kernels when that is
    // exact: one Mur extension with z faces only, float4 lanes (a face's
    // line and shift cells then belong to one thread), in-place steps, no
    // UPML, and no ADE cell on a face's line or shift plane (ADE post runs
    // between the E update and the Mur post step).
    MurZ mz{};
    bool fold_mur = e->fold_mur && !host && murs.size() == 1 && !upml && !fused && tl.vec == 4 && e->g.nz >= 4;
    if (fold_mur)
        for (const MurDev& m : murs[0]->mur_faces) {
            if (m.ny != 2) {
                fold_mur = false;
                break;
            }
            const int side = m.line == 0 ? 0 : 1;
            mz.on[side] = 1;
            mz.coeff[side] = m.coeff;
        }
    if (fold_mur)
        for (auto* x : ades) {
            if (x->ade_kmax < x->ade_kmin) continue;  // no cells
            if (mz.on[0] && x->ade_kmin <= 1) fold_mur = false;
            if (mz.on[1] && x->ade_kmax >= e->g.nz - 2) fold_mur = false;
        }
    if (!fold_mur) mz = MurZ{};
    // Dispersive extensions whose cells are disjoint share one pre and one
    // post launch per step (no cell is updated by two of them, so their list
    // order does not matter).
    bool merge_ade = e->merge_ade && ades.size() >= 2 && ades.size() <= (size_t)MAX_MERGE_ADE;
    if (merge_ade) {
        std::vector<std::pair<uint32_t, int>> all;
        for (size_t q = 0; q < ades.size(); ++q)
            for (uint32_t c : ades[q]->ade_cells) all.push_back({c, (int)q});
        std::sort(all.begin(), all.end());
        for (size_t q = 1; q < all.size() && merge_ade; ++q)
            if (all[q].first == all[q - 1].first && all[q].second != all[q - 1].second) merge_ade = false;
    }
    AdeMulti ade_pre_m{}, ade_post_m{};
    if (merge_ade)
        for (auto* x : ades) {
            if (x->ade.mode == OE_MODE_PHYSICAL) ade_pre_m.a[ade_pre_m.n++] = x->ade;
            ade_post_m.a[ade_post_m.n++] = x->ade;
        }
    for (const Launch& L : plan)
        if (L.a.t.vec != 4) fold_mur = false;  // the folded Mur step needs float4 lanes
    if (!fold_mur) mz = MurZ{};
    std::vector<Launch> plan_e = plan;  // the E launches carry the folded Mur faces
    for (Launch& L : plan_e) L.a.mur_z = mz;
    // Two-half-step batches advance the device step counter in their first H
    // launch (from -1, so that step st reads st); fused batches launch a
    // counter kernel at the end of each step (from 0).
    std::vector<Launch> plan_h = plan;
    const bool step_in_h = !fused && !plan_h.empty();
    if (step_in_h) plan_h[0].a.step_inc = step_dev;
    HIP_OK(hipMemsetAsync(step_dev, step_in_h ? 0xff : 0, sizeof(uint64_t), e->stream));

    // The slab launches of a half-step write other cells than the interior
    // launch and read only the other field (and their own psi), so they run
    // on the side stream beside it: fork after the previous work, join
    // before the next half-step.  Steps whose launches are timed stay on
    // one stream.
    bool overlap = e->overlap_slabs && !fused && !upml && !host;
    {
        bool any_slab = false, any_inner = false;
        for (const Launch& L : plan) ((L.axes & (AX_X | AX_Y)) ? any_slab : any_inner) = true;
        overlap = overlap && any_slab && any_inner;
    }
    if (overlap && !e->side) {
        HIP_OK(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming));
    }
    auto overlapped_half = [&](bool h, uint64_t st) -> oe_status {
        auto one = [&](const Launch& L, hipStream_t s) -> hipError_t {
            if (h) return launch_update_h(L.a, cp, L.axes, s);
            if (!fold_sources) return launch_update_e(L.a, cp, L.axes, s);
            YeeArgs a = L.a;
            a.exc = e->exc_table.as<PointExc>();
            a.n_exc = (int)px.size();
            for (size_t q = 0; q < px.size(); ++q) a.exc_idx[q] = px[q].idx;
            a.exc_step = st;
            a.exc_step_dev = step_dev;
            return launch_update_e(a, cp, L.axes, s);
        };
        const std::vector<Launch>& pl = h ? plan_h : plan_e;
        HIP_OK(hipEventRecord(e->ev_fork, e->stream));
        HIP_OK(hipStreamWaitEvent(e->side, e->ev_fork, 0));
        for (const Launch& L : pl)
            if (L.axes & (AX_X | AX_Y)) HIP_OK(one(L, e->side));
        for (const Launch& L : pl)
            if (!(L.axes & (AX_X | AX_Y))) HIP_OK(one(L, e->stream));
        HIP_OK(hipEventRecord(e->ev_join, e->side));
        HIP_OK(hipStreamWaitEvent(e->stream, e->ev_join, 0));
        return OE_OK;
    };

    // Host-callback hooks at point `which` (0..5): fields to the host, every
    // selected hook in list order, then the part the hook point may change
    // back to HBM (H after 0/1, E after 2/3, all after 4/5).
    auto run_hooks = [&](int which) -> oe_status {
        bool any = false;
        for (auto* x : hosts) any = any || ((x->host.hook_mask >> which) & 1u);
        if (!any) return OE_OK;
        const size_t N = e->g.N;
        if (!e->host_fields) HIP_OK(hipHostMalloc((void**)&e->host_fields, 6 * N * sizeof(float), hipHostMallocDefault));
        HIP_OK(hipMemcpyAsync(e->host_fields, e->F, 6 * N * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        HIP_OK(hipStreamSynchronize(e->stream));
        float* f[6];
        for (int d = 0; d < 6; ++d) f[d] = e->host_fields + (size_t)d * N;
        for (auto* x : hosts) {
            if (!((x->host.hook_mask >> which) & 1u)) continue;
            const int rc = x->host.hook(x->host.user, which, e->timestep, f);
            if (rc)
                return fail(rc >= 1 && rc <= 7 ? (oe_status)rc : OE_ERR_NUMERICAL,
                            "host extension hook %d returned %d at timestep %llu", which, rc,
                            (unsigned long long)e->timestep);
        }
        const size_t lo = which == 0 || which == 1 ? 3 : 0, cnt = which >= 4 ? 6 : 3;
        HIP_OK(hipMemcpyAsync(e->F + lo * N, e->host_fields + lo * N, cnt * N * sizeof(float), hipMemcpyHostToDevice,
                              e->stream));
        return OE_OK;
    };

    auto enqueue_step = [&](uint64_t st, uint64_t cur) -> oe_status {
        const bool ovl = overlap && !(e->profiling && e->prof_step);
        if (fused) {
            // slab H (old -> new buffer), the fused interior, slab E; then the
            // new buffer becomes the current one
            CpmlArgs cf{};
            if (cp) {
                cf = *cp;
                for (int f = 4; f < 6; ++f)
                    if (cf.on_mask & (1 << f)) {
                        cf.phw_pos[f] = cpx->zalt_pos[f - 4];
                        cf.phw_neg[f] = cpx->zalt_neg[f - 4];
                    }
            }
            float* Fo = e->F;
            float* Fn = e->F2;
            for (const Launch& L : slabs) {
                YeeArgs a = L.a;
                a.Fin = Fo;
                a.Fout = Fn;
                Timed tm(e, "update_h:slabs");
                HIP_OK(launch_update_h(a, &cf, L.axes, e->stream));
            }
            {
                fa.Fo = Fo;
                fa.Fn = Fn;
                Timed tm(e, cp ? "fused:interior" : "fused:grid");
                HIP_OK(launch_fused(fa, cp ? &cf : nullptr, e->stream));
            }
            for (const Launch& L : slabs) {
                YeeArgs a = L.a;
                a.Fin = Fo;
                a.Fout = Fn;
                Timed tm(e, "update_e:slabs");
                HIP_OK(launch_update_e(a, &cf, L.axes, e->stream));
            }
            std::swap(e->F, e->F2);
            if (cpx)
                for (int f = 4; f < 6; ++f)
                    if (cpx->cpml.on_mask & (1 << f)) {
                        std::swap(cpx->cpml.ph_pos[f], cpx->zalt_pos[f - 4]);
                        std::swap(cpx->cpml.ph_neg[f], cpx->zalt_neg[f - 4]);
                        cpx->cpml.phw_pos[f] = cpx->cpml.ph_pos[f];
                        cpx->cpml.phw_neg[f] = cpx->cpml.ph_neg[f];
                    }
        }
        if (upml) {
            Timed tm(e, "upml");
            HIP_OK(launch_upml_pre(e->F, e->g, upml->upml, false, e->stream));
        }
        if (host) OE_TRY(run_hooks(0));
        if (!fused && ovl) {
            OE_TRY(overlapped_half(true, st));
        } else if (!fused)
            for (const Launch& L : plan_h) {
                Timed tm(e, (L.axes & (AX_X | AX_Y)) ? "update_h:slabs" : (cp ? "update_h:interior" : "update_h:grid"));
                HIP_OK(launch_update_h(L.a, cp, L.axes, e->stream));
            }
        if (upml) {
            Timed tm(e, "upml");
            HIP_OK(launch_upml_post(e->F, e->g, upml->upml, false, e->stream));
        }
        if (host) OE_TRY(run_hooks(1));
        for (auto* x : murs) {
            if (fold_mur) break;
            Timed tm(e, "mur_pre");
            HIP_OK(launch_mur_pre(e->F, e->g, x->mur_dev.as<MurDev>(), (int)x->mur_faces.size(), e->stream));
        }
        if (merge_ade) {
            if (ade_pre_m.n) {
                Timed tm(e, "ade_pre");
                HIP_OK(launch_ade_multi(e->F, e->Cf, e->g.N, ade_pre_m, true, e->stream));
            }
        } else {
            for (auto* x : ades)
                if (x->ade.mode == OE_MODE_PHYSICAL) {
                    Timed tm(e, "ade_pre");
                    HIP_OK(launch_ade_pre(e->F, e->g.N, x->ade, e->stream));
                }
        }
        if (upml) {
            Timed tm(e, "upml");
            HIP_OK(launch_upml_pre(e->F, e->g, upml->upml, true, e->stream));
        }
        if (host) OE_TRY(run_hooks(2));
        if (!fused && ovl) {
            OE_TRY(overlapped_half(false, st));
        } else if (!fused)
            for (const Launch& L : plan_e) {
                Timed tm(e, (L.axes & (AX_X | AX_Y)) ? "update_e:slabs" : (cp ? "update_e:interior" : "update_e:grid"));
                if (fold_sources) {
                    YeeArgs a = L.a;
                    a.exc = e->exc_table.as<PointExc>();
                    a.n_exc = (int)px.size();
                    for (size_t q = 0; q < px.size(); ++q) a.exc_idx[q] = px[q].idx;
                    a.exc_step = st;
                    a.exc_step_dev = step_dev;
                    HIP_OK(launch_update_e(a, cp, L.axes, e->stream));
                } else {
                    HIP_OK(launch_update_e(L.a, cp, L.axes, e->stream));
                }
            }
        if (upml) {
            Timed tm(e, "upml");
            HIP_OK(launch_upml_post(e->F, e->g, upml->upml, true, e->stream));
        }
        if (merge_ade) {
            Timed tm(e, "ade_post");
            HIP_OK(launch_ade_multi(e->F, e->Cf, e->g.N, ade_post_m, false, e->stream));
        } else {
            for (auto* x : ades) {
                Timed tm(e, "ade_post");
                HIP_OK(launch_ade_post(e->F, e->Cf, e->g.N, x->ade, e->stream));
            }
        }
        for (auto* x : murs) {
            if (fold_mur) break;
            Timed tm(e, "mur_post");
            HIP_OK(launch_mur_post(e->F, e->g, x->mur_dev.as<MurDev>(), (int)x->mur_faces.size(), e->stream));
        }
        if (host) OE_TRY(run_hooks(3));
        if (point_ops) {
            Timed tm(e, "point_ops");
            HIP_OK(launch_point_ops(e->F, e->g.N, e->exc_table.as<PointExc>(), fold_sources ? 0 : (int)px.size(), st,
                                    e->probe_table.as<PointProbe>(), (int)pp.size(), e->stream, step_dev));
        }
        for (size_t q = 0; q < dfts.size(); ++q) {
            oe_extension* x = dfts[q];
            if (dev_dft) {
                Timed tm(e, "dft");
                HIP_OK(launch_dft_dev(e->F, e->g.N, x->dft, step_dev, dft0[q], (double)t_batch0, e->dt, e->stream));
            } else {
                if (x->dft_seen++ % x->dft_stride != 0) continue;
                Timed tm(e, "dft");
                HIP_OK(launch_dft(e->F, e->g.N, x->dft, x->dft_count, (double)(cur + 1) * e->dt,
                                  ((double)cur + 0.5) * e->dt, e->stream));
                x->dft_count += 1;
            }
        }
        if (!step_in_h) HIP_OK(launch_sweep_step(step_dev, e->stream));
        return OE_OK;
    };

    const int LG = e->graph_steps;
    const bool use_graph = LG > 0 && !fused && dev_dft && !host;
    struct GraphGuard {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        ~GraphGuard() {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
        }
    } graph;
    auto capture = [&]() -> oe_status {
        const bool ps = e->prof_step;
        e->prof_step = false;
        HIP_OK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        oe_status cs = OE_OK;
        for (int q = 0; q < LG && cs == OE_OK; ++q) cs = enqueue_step(0, 0);
        const hipError_t ce = hipStreamEndCapture(e->stream, &graph.g);
        e->prof_step = ps;
        if (cs != OE_OK) return cs;
        HIP_OK(ce);
        HIP_OK(hipGraphInstantiate(&graph.x, graph.g, nullptr, nullptr, 0));
        return OE_OK;
    };

    if (host) OE_TRY(run_hooks(4));  // Extension::pre_batch
    uint64_t st = 0;
    while (st < num_steps) {
        const uint64_t cur = e->timestep;
        // steps from cur on after which the host has nothing to do
        uint64_t nfree = num_steps - st;
        if (b->sample_interval > 0)
            nfree = std::min<uint64_t>(nfree, (b->sample_interval - cur % b->sample_interval) % b->sample_interval);
        if (need_decay)
            nfree = std::min<uint64_t>(nfree, cur == 0 ? check_interval : (check_interval - cur % check_interval) % check_interval);
        if (e->profiling)
            nfree = std::min<uint64_t>(nfree, ((uint64_t)e->profiling - st % (uint64_t)e->profiling) % (uint64_t)e->profiling);
        if (use_graph && nfree >= (uint64_t)LG) {
            if (!graph.x) OE_TRY(capture());
            HIP_OK(hipGraphLaunch(graph.x, e->stream));
            e->timestep += LG;
            executed += LG;
            st += LG;
            for (auto* x : dfts) {
                x->dft_seen += LG;
                x->dft_count += LG;
            }
            continue;
        }
        e->prof_step = e->profiling > 0 && (st % (uint64_t)e->profiling) == 0;
        OE_TRY(enqueue_step(st, cur));
        if (dev_dft)
            for (auto* x : dfts) {
                x->dft_seen += 1;
                x->dft_count += 1;
            }
        e->timestep += 1;
        ++executed;
        ++st;

        if (b->sample_interval > 0 && cur % b->sample_interval == 0) {
            if (sample_steps.size() >= e->energy_dev_cap) OE_TRY(drain());
            Timed tm(e, "energy");
            HIP_OK(launch_energy(e->F, e->g.N, e->partials, energy_blocks(), e->energy_dev + 2 * sample_steps.size(), e->stream));
            sample_steps.push_back(cur);
        }
        if (cur > 0 && cur % check_interval == 0 && (need_decay || host)) {
            // Extension::check_termination first; the decay test takes
            // precedence when it ends the batch at the same step
            // (src/fdtd/basic_engine.rs:115-147).
            bool ext_stop = false;
            for (auto* x : hosts) {
                if (!x->host.check_termination) continue;
                char why[256] = {0};
                if (x->host.check_termination(x->host.user, why, sizeof(why))) {
                    why[sizeof(why) - 1] = 0;
                    x->host_reason = why;
                    ext_stop = true;
                    break;
                }
            }
            if (need_decay) {
                const bool sampled_now = !sample_steps.empty() && sample_steps.back() == cur;
                OE_TRY(drain());
                double curE = last_total;
                if (!sampled_now) {  // compute_energy_sample at the check step
                    double ee, he;
                    OE_TRY(oe_engine_energy(e, &ee, &he));
                    curE = ee + he;
                }
                if (peak > 0.0 && curE > 0.0) {  // src/fdtd/basic_engine.rs:128-141
                    const double ddb = 10.0 * std::log10(curE / peak);
                    if (ddb < -std::fabs(b->energy_decay_db)) {
                        reason = OE_TERM_ENERGY_DECAY;
                        final_db = ddb;
                        break;
                    }
                }
            }
            if (ext_stop) {
                reason = OE_TERM_EXTENSION_STOP;
                break;
            }
        }
        if (e->profiling && e->pending.size() > 4096) {
            HIP_OK(hipStreamSynchronize(e->stream));
            collect_timing(e);
        }
    }
    OE_TRY(drain());
    HIP_OK(hipStreamSynchronize(e->stream));
    if (e->profiling) collect_timing(e);
    if (b->persistent_peak) e->peak_persistent = peak;
    if (host) {  // Extension::post_batch
        OE_TRY(run_hooks(5));
        HIP_OK(hipStreamSynchronize(e->stream));
    }

    r->timesteps_executed = executed;
    r->termination_reason = reason;
    r->final_decay_db = final_db;
    r->n_energy_samples = n_samples_total;
    r->elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return OE_OK;
}

}  // extern "C"
