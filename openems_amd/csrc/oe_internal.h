// oe_internal.h -- types shared by the host runtime (oe_api.hip) and the
// gfx950 kernels (oe_yee.hip, oe_aux.hip, oe_nf2ff.hip).  Not part of the
// public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

namespace oe {

constexpr double C0 = 299792458.0;
constexpr double EPS0 = 8.854187817e-12;
constexpr double MU0 = 1.256637062e-6;
constexpr double Z0 = 376.730313668;

// Owned device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = n;
        return n ? hipMalloc(&p, n) : hipSuccess;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Field block: one allocation of 6 arrays of N floats, E(x,y,z) then H(x,y,z).
// Coefficient block: 12 arrays of N floats, Ca(x,y,z) Cb(x,y,z) Da(x,y,z) Db(x,y,z).
struct Grid {
    int nx, ny, nz;
    size_t N;
};

// Yee kernel launch shape.  A work-group covers bx*V consecutive k of `by`
// rows (j) and marches along i over `ci` planes, keeping the i+1 (H update)
// or i-1 (E update) plane of the marched components in registers.
struct Tiles {
    int vec;  // 4: float4 accesses (nz % 4 == 0), 1: scalar
    int bx;   // lanes along k (power of two)
    int by;   // rows along j; bx * by == 256
    int ci;   // planes along i per work-group
};

// CPML state as the fused Yee kernels see it.  Face f = 2*dir + (max side).
// psi layout for face f: dims (nx,ny,nz) with axis `dir` replaced by L[f],
// z fastest.  tab holds, per face, 6 tables of Lmax floats:
// b_e, b_h, a_e, a_h, (1/kappa_e - 1), (1/kappa_h - 1).
constexpr int CPML_TABLES = 6;
enum { TAB_BE = 0, TAB_BH = 1, TAB_AE = 2, TAB_AH = 3, TAB_KE = 4, TAB_KH = 5 };

struct CpmlArgs {
    int on_mask;   // bit f set when face f is active
    int physical;  // 1 = apply the field correction
    int Lmax;      // table stride, a multiple of 4
    int zvec;      // z faces can be handled 4 cells at a time (L % 4 == 0, nz % 4 == 0)
    int L[6];
    int g0[6];     // first global index of the slab along its axis
    float* pe_pos[6];
    float* pe_neg[6];
    float* ph_pos[6];
    float* ph_neg[6];
    // where the H updates store the new H psi: the read arrays (in place)
    // except in a ping-pong step, whose z faces read the old and write the
    // new buffer
    float* phw_pos[6];
    float* phw_neg[6];
    const float* tab;  // [6][CPML_TABLES][Lmax]
    const float* dummy;  // 16 zeroed floats: load target for cells outside every slab
};

// A launch covers up to 4 boxes [i0,i1) x [j0,j1) x [0,nz); work-group ids
// are split over the boxes by prefix sums (first[b] .. first[b+1]).
constexpr int MAX_BOXES = 4;
struct Box {
    int i0, i1, j0, j1;
    int n_jb, n_ib;  // j-blocks, i-chunks of this box
};

struct PointExc;
constexpr int MAX_FOLD_SOURCES = 8;

// Mur faces folded into the E update: a z face's line and shift cells are
// lanes of one float4 lane group, so the thread that updates them also takes
// the pre value, the post correction and the boundary write.
struct MurZ {
    int on[2];       // [k = 0 face, k = nz - 1 face]
    float coeff[2];
};

struct YeeArgs {
    Grid g;
    // Fields read (Fin) and written (Fout); equal for an in-place update.
    // The H half reads E and the old H from Fin and writes H to Fout; the E
    // half reads the new H from Fout and the old E from Fin and, when the
    // two differ, writes every E cell (unchanged ones copied) to Fout.
    const float* Fin;
    float* Fout;
    const float* Cf; // coefficients
    // Material table (oe_coef.hip): per-cell row indices and rows of
    // (a_x, b_x, a_y, b_y, a_z, b_z) for the H (Da, Db) and E (Ca, Cb)
    // half-steps; a half whose row count is 0 reads the coefficient arrays.
    const uint16_t* mat_h;
    const uint16_t* mat_e;
    // the same indices one byte each, for a half of 2 to 256 rows (else null)
    const uint8_t* mat8_h;
    const uint8_t* mat8_e;
    const float* tab_h;
    const float* tab_e;
    int n_mat_h, n_mat_e;
    Tiles t;
    int n_kb;              // blocks along k (every box spans all k)
    int nbox;
    Box box[MAX_BOXES];
    int first[MAX_BOXES + 1];
    int xcd_remap;         // 1: make consecutive tiles share an XCD (L2)
    // Point sources added by the E half-step after its update (when no
    // extension acts between the E update and the sources): the lane holding
    // a source's cell adds or sets samples[exc_step], in list order.  The
    // cell indices travel in the kernel arguments, so the per-plane test is
    // a comparison with uniform values.
    const PointExc* exc;
    int n_exc;
    uint32_t exc_idx[MAX_FOLD_SOURCES];
    uint64_t exc_step;
    const uint64_t* exc_step_dev;  // non-null: the step comes from the device counter
    // First-order Mur on the z faces applied by the E half-step itself (see
    // e_body): on[0] the k = 0 face, on[1] the k = nz - 1 face.
    MurZ mur_z;
    // Non-null on one H launch per step: its first lane advances the device
    // step counter before anything of the step reads it (the counter starts
    // a batch at -1).
    uint64_t* step_inc;
};

// CPML code compiled into a launch: bit 0 x faces, bit 1 y faces, bit 2 z faces.
enum { AX_X = 1, AX_Y = 2, AX_Z = 4 };

// Largest material table a half-step keeps in LDS (rows of 24 bytes).
constexpr int MAX_MAT_ROWS = 1024;
// Builds the material table of the H (h) or E half from the coefficient
// block (oe_coef.hip); *n_rows = 0 when the grid holds more than cap
// distinct coefficient sets.
hipError_t build_material_table(const float* Cf, size_t N, bool h, int cap, uint16_t* mat, float* tab,
                                int* n_rows, hipStream_t s);
// Copies n 16-bit row indices of a table of at most 256 rows to bytes.
hipError_t narrow_material_index(const uint16_t* in, uint8_t* out, size_t n, hipStream_t s);

// Fused step (oe_yee.hip): one pass over a box computes H^{n+1/2} and
// E^{n+1} of every cell, reading the old fields from Fo and writing the new
// ones to Fn.  A work-group holds FUSED_ROWS rows of full k-rows (one wave
// per row, nz <= 256) plus one wave that recomputes the H of the row below
// the block; the H of the plane before its i-chunk is recomputed too.  Rows
// and planes outside the box (CPML x/y slabs) must already hold their new H
// in Fn.  The z-face H psi is read from c.ph_* and written to c.phw_*.
constexpr int FUSED_ROWS = 4;
struct FusedArgs {
    Grid g;
    const float* Fo;
    float* Fn;
    const float* Cf;
    const uint16_t* mat_h;
    const uint16_t* mat_e;
    const float* tab_h;
    const float* tab_e;
    int n_mat_h, n_mat_e;  // both > 0, or both 0 (coefficient arrays)
    int i0, i1, j0, j1;    // box
    int ci;                // planes per work-group
    int n_jb, n_ib;
    int xcd_remap;
    const uint8_t* mat8_e;  // byte-wide E indices (2 to 256 rows) or null
    int kind;               // 1: rows exchanged through LDS; 2: recomputed per wave (k_fused_rows)
};
// Whether the fused step handles this grid (V = 4 lanes, every k-row inside
// one wave).  Below nz = 128 most lanes of a row idle, so by default the
// engine fuses only from there on.
bool fused_supported(const Grid& g);
hipError_t launch_fused(const FusedArgs& a, const CpmlArgs* c, hipStream_t s);

// Kernel launchers (oe_yee.hip).
hipError_t launch_update_h(const YeeArgs& a, const CpmlArgs* c, int axes, hipStream_t s);
hipError_t launch_update_e(const YeeArgs& a, const CpmlArgs* c, int axes, hipStream_t s);

// Sweep: many structures per launch.  A structure is its grid, field and
// coefficient blocks and CPML state; a box is a part of one structure with
// that structure's tile shape; wg_box maps every work-group to its box.
struct SweepStruct {
    Grid g;
    float* F;
    const float* Cf;
    CpmlArgs c;
    // the structure's material table (see YeeArgs); used by launches whose
    // every structure has one
    const uint16_t* mat_h;
    const uint16_t* mat_e;
    const uint8_t* mat8_h;  // byte-wide indices (2 to 256 rows) or null
    const uint8_t* mat8_e;
    const float* tab_h;
    const float* tab_e;
    int n_mat_h, n_mat_e;
};
struct SweepBox {
    int sid;
    int i0, i1, j0, j1;
    int n_jb, n_kb;
    int bx, by, ci;
    int ns;     // i-chunks of ci planes per work-group (boxes with fewer j rows than 256 / bx)
    int first;  // first work-group of the box
};
struct SweepArgs {
    const SweepStruct* st;
    const SweepBox* box;
    const int* wg_box;
    int xcd_remap;
    uint64_t* step_inc;  // non-null on the step's first H launch: lane 0 advances the step counter
};
// h: H (else E) half-step over n_wg work-groups; axes 0, AX_Z, AX_X|AX_Z,
// AX_Y|AX_Z or all faces.  rows > 0: every structure reads its material
// table, rows being the largest table of the launch's half (LDS size).
hipError_t launch_sweep(bool h, const SweepArgs& a, int n_wg, int axes, int vec, int rows, hipStream_t s);

// Auxiliary kernels (oe_aux.hip).
hipError_t launch_fill(float* p, size_t n, float v, hipStream_t s);
// One region write of the coefficient block: cells [box0, box1) x [box2,
// box3) x [box4, box5) of array a get v[a] for every bit a of mask.
struct BoxFill {
    int box[6];
    unsigned mask;
    float v[12];
};
// Work-groups of one box in launch_boxes_fill.
int boxes_fill_blocks(const BoxFill& o);
// Pairwise-disjoint boxes in one launch; wg_first[o]: the first of box o's
// work-groups (the running sum of boxes_fill_blocks), n_wg their total.
hipError_t launch_boxes_fill(float* Cf, const Grid& g, const BoxFill* ops, const int* wg_first, int n_ops, int n_wg,
                             hipStream_t s);

struct PointExc {
    uint32_t idx;   // linear cell index
    uint32_t comp;  // 0..2 (E component)
    uint32_t soft;
    uint32_t pad;
    const float* samples;
    uint64_t n_samples;
};
struct PointProbe {
    uint32_t idx;
    uint32_t pad;
    double* out;    // [capacity][6]
    uint64_t capacity;
    uint64_t* count;  // device counter
};
// step_dev non-null: the step index is read from the device counter.
hipError_t launch_point_ops(float* F, size_t N, const PointExc* exc, int n_exc, uint64_t step,
                            const PointProbe* probes, int n_probes, hipStream_t s,
                            const uint64_t* step_dev = nullptr);

// Sum of squares of E (out[0]) and H (out[1]); partials has >= 2*blocks doubles.
hipError_t launch_energy(const float* F, size_t N, double* partials, int blocks, double* out,
                         hipStream_t s);
int energy_blocks();

// Mur ABC face (src/extensions/mur_abc.rs:59-151).
struct MurDev {
    int ny, nyp, nypp;
    int line, shift;
    int s0, s1;
    float coeff;
    float* v_nyp;
    float* v_nypp;
};
hipError_t launch_mur_pre(float* F, const Grid& g, const MurDev* faces, int n, hipStream_t s);
hipError_t launch_mur_post(float* F, const Grid& g, const MurDev* faces, int n, hipStream_t s);

// ADE (src/extensions/dispersive.rs).  One thread per listed cell.
struct AdeDev {
    int kind, mode;
    int n;
    const uint32_t* idx;  // linear indices
    float* J;             // [3][n]
    float* Jold;          // [3][n] (Lorentz physical: P^{n-1})
    float* Eprev;         // [3][n] (physical: E^n captured before the E update)
    float a, b, c, dt_t;  // compat coefficients (float, as the reference)
    float pa, pb, pc;     // physical coefficients
    double delta[3];      // cell sizes: K = Cb * delta per cell and component
    float dtf;
};
hipError_t launch_ade_pre(const float* F, size_t N, const AdeDev& a, hipStream_t s);
hipError_t launch_ade_post(float* F, const float* Cf, size_t N, const AdeDev& a, hipStream_t s);
// Up to MAX_MERGE_ADE extensions with disjoint cells in one launch (pre: the
// E^n capture, else the post-E step), passed by value.
constexpr int MAX_MERGE_ADE = 8;
struct AdeMulti {
    int n;
    AdeDev a[MAX_MERGE_ADE];
};
hipError_t launch_ade_multi(float* F, const float* Cf, size_t N, const AdeMulti& m, bool pre, hipStream_t s);

// UPML (src/extensions/pml.rs): up to 6 regions whose cells are numbered
// consecutively (region r holds cells first[r] .. first[r+1]); arrays are
// [8][3][total]: vv, vvfo, vvfn, ii, iifo, iifn, voltage flux, current flux.
struct UpmlDev {
    int n_reg;
    int start[6][3];
    int size[6][3];
    long first[7];
    float* a;
};
// e: the E (voltage) hooks, else the H (current) hooks.
hipError_t launch_upml_pre(float* F, const Grid& g, const UpmlDev& u, bool e, hipStream_t s);
hipError_t launch_upml_post(float* F, const Grid& g, const UpmlDev& u, bool e, hipStream_t s);

// A recording box's faces for the running DFT without term tables: face f
// holds the nodes (i1, i2) of the plane q of axis ax, i1 along
// a1 = (ax + 1) % 3 from lo1 (n1 nodes), i2 along a2 = (ax + 2) % 3 from lo2
// (n2 nodes), and per node the 4 signals E_a1, E_a2, H_a1, H_a2 from
// signal sig0 + 4 (node), node = (i1 - lo1) n2 + (i2 - lo2).
struct SurfFace {
    int ax, q, lo1, n1, lo2, n2, sig0;
};
constexpr int MAX_SURF_FACES = 6;
struct DftSurface {
    int n;  // faces (0: the signals come from the term tables)
    int nx, ny, nz;
    SurfFace f[MAX_SURF_FACES];
};

// Running DFT of weighted field signals (one thread per signal).
struct DftDev {
    int n_sig, n_f, prec;
    const uint64_t* off;
    const uint32_t* comp;
    const uint64_t* cell;
    const double* w;
    const uint8_t* is_h;
    const double* freqs;
    double* acc64;   // [n_f][n_sig][2]
    float* acc32;    // [n_f][n_sig][4]: sum re, sum im, compensation re, im
    double* series;  // [cap][n_sig] or null
    uint64_t series_cap;
    DftSurface surf;  // n > 0: signals formed from the box's indices, not the tables
};
hipError_t launch_dft(const float* F, size_t N, const DftDev& d, uint64_t sample, double tE, double tH,
                      hipStream_t s);
// Step from the device counter: sample = sample0 + st, times from t0 + st.
hipError_t launch_dft_dev(const float* F, size_t N, const DftDev& d, const uint64_t* step_dev, uint64_t sample0,
                          double t0, double dt, hipStream_t s);

// Sweep step pieces (oe_aux.hip).  The step index is read from *step.
struct SweepExc {
    float* target;           // field element
    const float* samples;
    uint64_t n_samples;
    uint32_t soft;
    uint32_t pad;
};
struct SweepDft {
    const float* F;          // the structure's field block
    size_t N;
    DftDev d;                // at most 256 signals
    uint64_t sample0;        // sample index of sweep step 0
    double t0;               // timestep of sweep step 0
};
// Excitations and port DFT of every active structure after its E update: one
// work-group per structure (first: excitation ranges, serial: 1 where the
// structure's source cells repeat, dft_of: index into dfts or -1).
hipError_t launch_sweep_point(const SweepExc* e, const int* first, const int* serial, const SweepDft* dfts,
                              const int* dft_of, int n_struct, const uint64_t* step, double dt, hipStream_t s);
hipError_t launch_sweep_step(uint64_t* step, hipStream_t s);
// Energies [n][2] of structures st[ids[0..n)] (partials: n * 2 * energy_blocks() doubles).
hipError_t launch_sweep_energy(const SweepStruct* st, const int* ids, int n, double* partials, double* out,
                               hipStream_t s);

// NF2FF (oe_nf2ff.hip).  One face of the closed surface: normal `axis`,
// outward sign, position c0 and node coordinates c1, c2 along the in-plane
// axes a1 = (axis+1)%3, a2 = (axis+2)%3.  Its phasors sit on the device as
// [n1][n2][4] complex values (E_a1, E_a2, H_a1, H_a2) in one of two layouts:
// src_kind 0 = float64 (re, im) pairs, 1 = float32 Kahan quads (sum re,
// sum im, compensation re, im); every value is multiplied by `scale`.
struct Nf2ffFace {
    int axis, sign;
    double c0;
    std::vector<double> c1, c2;
    const void* src;
    int src_kind;
    double scale;
};
struct Nf2ffOut {
    double* et_re;  // [n_theta][n_phi] each, or null
    double* et_im;
    double* ep_re;
    double* ep_im;
    double* et;     // [n_theta][n_phi][re, im] each, or null
    double* ep;
    double prad, dmax;
    double gemm_ms, total_ms, gemm_flop;
    double gemm_split_ms, gemm_split_flop;  // the part on the FP16 matrix cores (split operands)
};
// Angle plan (distinct direction cosines per axis) and work buffers of the
// transform, reused by later calls with the same angles.
struct Nf2ffCache;
Nf2ffCache* nf2ff_cache_new();
void nf2ff_cache_free(Nf2ffCache* c);
struct Nf2ffCacheDeleter {
    void operator()(Nf2ffCache* c) const { nf2ff_cache_free(c); }
};
typedef std::unique_ptr<Nf2ffCache, Nf2ffCacheDeleter> Nf2ffCachePtr;

// Far field at distance r for every (theta, phi) (matlab/AnalyzeNF2FF.m).
// Fails with hipErrorInvalidValue on empty input.
hipError_t nf2ff_compute(const std::vector<Nf2ffFace>& faces, double freq, const double* theta, int n_theta,
                         const double* phi, int n_phi, double r, const double center[3], hipStream_t s,
                         Nf2ffCache* cache, Nf2ffOut* out);

}  // namespace oe
