/*
 * oe.h -- C ABI of the MI355X (gfx950) FDTD timestep engine.
 *
 * Drop-in boundary for the timestep loop of the Rust openEMS port
 * (umgefahren/openems).  Every entry point names the reference interface it
 * replaces (paths relative to the reference root).  Plain C types only: host
 * pointers and sizes in, status codes out.  Field and coefficient arrays use
 * the reference's layout, one float32 array per component, z fastest:
 * index = (i * ny + j) * nz + k  (src/arrays/mod.rs:61-64).
 *
 * Threading: an engine handle is used by one caller at a time, as the
 * reference's `&mut self` methods require.  Each engine owns one HIP stream.
 * Errors never abort: every call returns an oe_status and oe_last_error()
 * gives the thread-local message (mirrors crate::Error, src/lib.rs:61-88).
 */
#ifndef OE_OE_H
#define OE_OE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
/* One code per variant of crate::Error (src/lib.rs:61-88), plus DEVICE for
 * HIP runtime failures, which the reference's GPU engine panics on
 * (src/fdtd/gpu_engine.rs:88,108). */
typedef enum {
    OE_OK = 0,
    OE_ERR_CONFIG = 1,
    OE_ERR_IO = 2,
    OE_ERR_GRID = 3,
    OE_ERR_MATERIAL = 4,
    OE_ERR_NUMERICAL = 5,
    OE_ERR_XML = 6,
    OE_ERR_DEVICE = 7
} oe_status;

/* Message of the last failed call on this thread ("" if none). */
const char* oe_last_error(void);

/* Library version string. */
const char* oe_version(void);

/* ---------------------------------------------------------------- operator */
typedef struct oe_operator oe_operator;

/* Grid::uniform + Operator::new (src/geometry/mod.rs:60-66,
 * src/fdtd/operator.rs:59-77): dt = 0.99 CFL, vacuum Ca=Da=1,
 * Cb = dt/(eps0 d), Db = -dt/(mu0 d).  delta[3] are the cell sizes in metres. */
oe_status oe_operator_create(size_t nx, size_t ny, size_t nz, const double delta[3],
                             oe_operator** out);

/* Grid::cartesian + Operator::new with explicit mesh lines (nx+1 .. lines per
 * axis).  As in the reference (src/geometry/mod.rs:113-118,
 * src/fdtd/operator.rs:82-89) only the minimum spacing per axis enters dt
 * and the coefficients. */
oe_status oe_operator_create_lines(const double* x, size_t nxl, const double* y, size_t nyl,
                                   const double* z, size_t nzl, oe_operator** out);

void oe_operator_destroy(oe_operator* op);

/* Operator::timestep (src/fdtd/operator.rs:145). */
double oe_operator_timestep(const oe_operator* op);

/* Operator::dimensions (src/fdtd/operator.rs:157). */
void oe_operator_dims(const oe_operator* op, size_t dims[3]);

/* Operator::set_material (src/fdtd/operator.rs:187-244).  region is
 * [i0, i1, j0, j1, k0, k1) half-open; later calls overwrite earlier ones. */
oe_status oe_operator_set_material(oe_operator* op, double eps_r, double mu_r, double sigma_e,
                                   double sigma_m, const size_t region[6]);

/* Operator::set_pec (src/fdtd/operator.rs:249-263): Ca = Cb = 0. */
oe_status oe_operator_set_pec(oe_operator* op, const size_t region[6]);

/* Lumped resistor R (ohm) on the `dir` edges of region, as an openEMS
 * LumpedElement (python/openEMS/ports.py:177-181; replaces the heuristic
 * LumpedRlc of src/extensions/lumped_rlc.rs:165-313).  Edges along dir are in
 * series, edge lines across it in parallel; each edge gets the conductivity
 * d_dir / (R_edge * A), R_edge = R * n_parallel / n_series, added to the
 * material under it (the last set_material covering the cell) before Ca/Cb
 * of component dir are formed.  PEC cells stay PEC. */
oe_status oe_operator_add_lumped_element(oe_operator* op, int dir, const size_t region[6], double R);

/* ------------------------------------------------------------------ engine */
typedef struct oe_engine oe_engine;
typedef struct oe_extension oe_extension;

/* EngineImpl::new (src/fdtd/engine_impl.rs:32; clone of the coefficient
 * arrays at src/fdtd/basic_engine.rs:42-54).  The operator's coefficients
 * are materialised in HBM; the operator is only borrowed. */
oe_status oe_engine_create(const oe_operator* op, int device, oe_engine** out);

/* EngineImpl::new from explicit coefficient arrays: coeff[12] are host
 * pointers to N floats each, ordered Ca_x, Ca_y, Ca_z, Cb_x, Cb_y, Cb_z,
 * Da_x, Da_y, Da_z, Db_x, Db_y, Db_z (the vv/vi/ii/iv arrays; Db keeps the
 * reference's negative sign, src/fdtd/operator.rs:129-131). */
oe_status oe_engine_create_from_arrays(size_t nx, size_t ny, size_t nz, const double delta[3],
                                       double dt, const float* const coeff[12], int device,
                                       oe_engine** out);

/* Frees the engine.  Extensions created on it stay valid handles but are
 * detached: every later call on them except oe_extension_destroy returns
 * OE_ERR_CONFIG. */
void oe_engine_destroy(oe_engine* eng);

/* EngineImpl::current_timestep (src/fdtd/engine_impl.rs:53). */
uint64_t oe_engine_timestep(const oe_engine* eng);

/* EngineImpl::read_fields (src/fdtd/engine_impl.rs:59): copies E then H
 * (6 arrays of N floats, each may be NULL to skip) from HBM. */
oe_status oe_engine_read_fields(oe_engine* eng, float* const fields[6]);

/* EngineImpl::write_fields (src/fdtd/engine_impl.rs:65): uploads the given
 * arrays (NULL entries are left unchanged). */
oe_status oe_engine_write_fields(oe_engine* eng, const float* const fields[6]);

/* Copy one of the 12 coefficient arrays back (order as in
 * oe_engine_create_from_arrays). */
oe_status oe_engine_read_coefficients(oe_engine* eng, int which, float* out);

/* Coefficient layout of the Yee updates; no counterpart in the reference,
 * whose engines read the 12 arrays cloned from the Operator
 * (src/fdtd/basic_engine.rs:42-54).  By default (unless the environment sets
 * OE_COEF_TABLE=0) an engine derives from its coefficient arrays a material
 * table: a 2-byte row index per cell and rows holding each distinct set of
 * (a_x, b_x, a_y, b_y, a_z, b_z), separately for the H half-step (Da, Db)
 * and the E half-step (Ca, Cb).  The rows are the arrays' own float values,
 * so results are unchanged; a half-step then reads 2 bytes of coefficients
 * per cell instead of 24.  A half whose grid holds more than 1024 distinct
 * sets keeps reading the arrays.  enable: 1 builds the table if absent, 0
 * drops it (every half reads the arrays), -1 only queries.  rows[0], rows[1]
 * (may be NULL) receive the E and H row counts, 0 for a half that reads the
 * arrays. */
oe_status oe_engine_material_table(oe_engine* eng, int enable, int rows[2]);

/* Step scheme of oe_engine_run_batch; no counterpart in the reference, whose
 * engines run the H and E half-steps as two sweeps
 * (src/fdtd/basic_engine.rs:92-95).  Mode 0 (the default) runs the two
 * half-steps in place.  With mode 1 a batch whose extensions add nothing
 * between the half-steps (CPML, probes, DFT and ports only; no Mur, ADE or
 * UPML) on a grid with nz a multiple of 4 between 128 and 256 runs one fused
 * pass per timestep over the interior of the CPML x/y slabs: the pass reads
 * the old fields from one buffer and writes the new ones to a second, so the
 * engine holds two field blocks once a fused batch has run.  Every cell's
 * arithmetic is unchanged, so results are identical.  Mode 2 fuses from
 * nz = 4 on.  On MI355X the fused pass moves fewer bytes but runs slower
 * than the two half-steps with CPML (its per-plane barrier and register
 * count limit the loads in flight), hence the default.  The environment
 * variable OE_FUSED sets the initial mode.  -1 only queries; *current (may
 * be NULL) receives the mode in force. */
oe_status oe_engine_fused_steps(oe_engine* eng, int mode, int* current);

/* Launch scheme of oe_engine_run_batch; no counterpart in the reference,
 * whose GPU engine submits every step from the host
 * (src/fdtd/gpu_engine.rs:1039-1271).  The kernels of a step read the step
 * index from a device counter, so a run of `steps` consecutive steps after
 * which the host has nothing to do (no energy sample, decay check or timed
 * step falls inside) is captured once per batch into a HIP graph and
 * replayed; other steps are launched one by one.  Results are identical
 * either way.  steps = 0 launches every step directly; the default is 16
 * (environment variable OE_GRAPH_STEPS); -1 only queries.  Batches with a
 * DFT sampled every n > 1 steps, and fused steps, launch directly.
 * *current (may be NULL) receives the value in force. */
oe_status oe_engine_graph_steps(oe_engine* eng, int steps, int* current);

/* EngineImpl::reset (src/fdtd/engine_impl.rs:70): zero fields and the step
 * counter (extension state is reset with oe_extension_reset). */
oe_status oe_engine_reset(oe_engine* eng);

/* Sum of squares of E and of H in float64 (Field3D::energy,
 * src/arrays/field.rs:219-221; VectorField3D::energy :314-316). */
oe_status oe_engine_energy(oe_engine* eng, double* e_energy, double* h_energy);

/* -------------------------------------------------------------- batch ABI */
/* ScheduledExcitation with a Sampled waveform (src/fdtd/batch.rs:71-100):
 * samples[s] is applied after step s of the batch (soft: E += v, hard:
 * E = v), exactly as BasicEngine::apply_excitations
 * (src/fdtd/basic_engine.rs:313-395).  Steps past n_samples apply nothing. */
typedef struct {
    uint32_t i, j, k;
    uint32_t direction; /* 0 = x, 1 = y, 2 = z */
    uint32_t soft;      /* 1 = soft (add), 0 = hard (set) */
    const float* samples;
    uint64_t n_samples;
} oe_excitation;

/* EnergySample (src/fdtd/batch.rs:197-209). */
typedef struct {
    uint64_t timestep;
    double e_energy;
    double h_energy;
    double total_energy;
} oe_energy_sample;

/* EngineBatch + TerminationConfig + EnergyMonitorConfig
 * (src/fdtd/batch.rs:17-32, 150-193).  has_* flags encode Option<T>. */
typedef struct {
    int has_num_steps;
    uint64_t num_steps;
    const oe_excitation* excitations;
    size_t n_excitations;
    oe_extension* const* extensions;
    size_t n_extensions;
    /* TerminationConfig */
    int has_max_timesteps;
    uint64_t max_timesteps;
    int has_energy_decay_db;
    double energy_decay_db;
    uint64_t check_interval; /* reference default 100 */
    /* EnergyMonitorConfig */
    uint64_t sample_interval; /* 0 = disabled */
    int track_peak;
    /* Extension to the reference (fixes its per-batch peak reset): when set,
     * the peak used by the decay test persists across batches until
     * oe_engine_reset. */
    int persistent_peak;
    /* Caller-owned output buffer for energy samples. */
    oe_energy_sample* energy_samples;
    size_t energy_samples_capacity;
} oe_batch_desc;

/* TerminationReason (src/fdtd/batch.rs:52-67). */
typedef enum {
    OE_TERM_STEPS_COMPLETED = 0,
    OE_TERM_ENERGY_DECAY = 1,
    OE_TERM_EXTENSION_STOP = 2
} oe_termination;

/* BatchResult (src/fdtd/batch.rs:36-48). */
typedef struct {
    uint64_t timesteps_executed;
    int termination_reason; /* oe_termination */
    double final_decay_db;
    size_t n_energy_samples; /* may exceed the capacity; only capacity are written */
    double elapsed_seconds;
} oe_batch_result;

/* EngineImpl::run_batch (src/fdtd/engine_impl.rs:48-50; loop of
 * BasicEngine::run_batch, src/fdtd/basic_engine.rs:56-163).  Per step:
 *   H update (+CPML H correction, fused) -> pre-E hooks (Mur capture, ADE
 *   capture) -> E update (+CPML E correction, fused) -> post-E hooks (ADE,
 *   lumped port, Mur) -> excitations -> probes/DFT -> energy sample and
 *   termination test. */
oe_status oe_engine_run_batch(oe_engine* eng, const oe_batch_desc* batch, oe_batch_result* result);

/* Kernel timing accumulated while profiling is on, measured with HIP events
 * on the engine's stream around every launch.  Names: "update_h:grid" and
 * "update_e:grid" (no CPML), or "update_h:interior" / "update_h:slabs" and
 * the "update_e:" pair (CPML: interior box with z-face CPML, x/y face slabs
 * with all faces), "mur_pre", "mur_post", "ade_pre", "ade_post",
 * "point_ops", "energy".  Returns the launch count and total milliseconds. */
oe_status oe_engine_kernel_time(oe_engine* eng, const char* name, uint64_t* launches, double* total_ms);

/* Per-launch event timing: 0 off (default); n > 0 records HIP events around
 * every launch of every n-th step of a batch (n = 1: every step). */
oe_status oe_engine_set_profiling(oe_engine* eng, int period);

/* Tile shape of the Yee kernels: bx lanes along z (in vector units), by rows
 * along y, ci planes along x per work-group; 0 keeps the automatic choice.
 * vec = 4 uses 16-byte accesses (requires nz % 4 == 0), 1 scalar, 0 auto. */
oe_status oe_engine_set_tiles(oe_engine* eng, int bx, int by, int ci, int vec);
oe_status oe_engine_get_tiles(const oe_engine* eng, int tiles[4]);

/* ------------------------------------------------------------- waveforms */
/* ExcitationType (src/fdtd/excitation.rs:9-42). */
typedef enum {
    OE_EXC_GAUSSIAN = 0,           /* exp(-((t - t0) / tau)^2) */
    OE_EXC_SINUSOIDAL = 1,         /* sin(2 pi f t) */
    OE_EXC_GAUSSIAN_MODULATED = 2, /* sin(2 pi f t) exp(-((t - t0) / tau)^2) */
    OE_EXC_DIRAC = 3,              /* 1 at |t| < 1e-15, else 0 */
    OE_EXC_STEP = 4,               /* 1 for t >= 0 */
    OE_EXC_CUSTOM = 5              /* linear interpolation of (times, values) */
} oe_waveform_type;

typedef struct {
    int type;         /* oe_waveform_type */
    double t0, tau;   /* Gaussian variants */
    double frequency; /* Sinusoidal, GaussianModulated */
    const double* times;  /* Custom */
    const double* values; /* Custom */
    size_t n_points;      /* Custom */
    double amplitude;
} oe_waveform;

/* Excitation::evaluate (src/fdtd/excitation.rs:150-200), amplitude included. */
oe_status oe_waveform_evaluate(const oe_waveform* w, double t, double* value);

/* ScheduledExcitation::from_excitation (src/fdtd/batch.rs:110-129):
 * out[s] = (float) evaluate((start_timestep + s) * dt) for s < n. */
oe_status oe_waveform_sample(const oe_waveform* w, uint64_t start_timestep, uint64_t n, double dt, float* out);

/* -------------------------------------------------------------- extensions */
/* Extensions are device-resident objects bound to one engine and passed to
 * run_batch by pointer (their state persists across batches).  They replace
 * the CPU callbacks of the Extension trait (src/extensions/extension_trait.rs:
 * 32-167) at the fixed hook points listed at oe_engine_run_batch. */

typedef enum { OE_MODE_COMPAT = 0, OE_MODE_PHYSICAL = 1 } oe_mode;

/* CpmlBoundaries + Cpml::new (src/extensions/cpml.rs:27-51, 235-304).
 * mode COMPAT reproduces the reference exactly (psi recursions updated, no
 * field correction, :343-421); PHYSICAL applies the Roden-Gedney correction
 * with the freshly updated psi.  Only one CPML may be attached per engine;
 * it is fused into the Yee kernels. */
typedef struct {
    size_t layers[6]; /* xmin, xmax, ymin, ymax, zmin, zmax */
    double grading_order;    /* default 3 */
    double reflection_coeff; /* default 1e-6 */
    double alpha_max;        /* default 0 */
    double kappa_max;        /* default 1 */
    int mode;                /* oe_mode */
} oe_cpml_config;
oe_status oe_ext_cpml_create(oe_engine* eng, const oe_cpml_config* cfg, oe_extension** out);

/* MurAbcConfig + MurAbc::new (src/extensions/mur_abc.rs:14-28, 173-200).
 * phase_velocity <= 0 selects c0. */
oe_status oe_ext_mur_create(oe_engine* eng, const int enabled[6], double phase_velocity,
                            oe_extension** out);

/* LorentzMaterial / DrudeMaterial / DebyeMaterial ::new
 * (src/extensions/dispersive.rs:134-159, 476-498, 570-592).
 * params: Lorentz {omega_0, gamma, delta_eps}; Drude {omega_p, gamma};
 * Debye {tau, delta_eps}.  positions: n triples (i, j, k). */
typedef enum { OE_ADE_LORENTZ = 0, OE_ADE_DRUDE = 1, OE_ADE_DEBYE = 2 } oe_ade_kind;
oe_status oe_ext_ade_create(oe_engine* eng, int kind, const double params[3],
                            const uint32_t* positions, size_t n, int mode, oe_extension** out);

/* Upml (src/extensions/pml.rs:209-667) with PmlBoundaries {layers,
 * grading_order, reflection_coeff}: up to six non-overlapping regions (x
 * faces over the whole y-z plane, y faces without the x regions, z faces
 * without both), per cell and component the coefficients vv, vvfo, vvfn
 * (E) and ii, iifo, iifn (H) of Taflove eq. 7.88/7.90 as the reference forms
 * them, and a voltage and a current flux.  Hooks: before the H update
 * f = ii H - iifo flux, H <- flux, flux <- f; after it flux <- H,
 * H <- f + iifn H; the same with the v-coefficients around the E update
 * (pre/post_current_update, pre/post_voltage_update, :500-657).  This is the
 * reference's formulation exactly, including its instability (its own tests
 * are disabled for it); CPML is the working absorber. */
oe_status oe_ext_upml_create(oe_engine* eng, const size_t layers[6], double grading_order,
                             double reflection_coeff, oe_extension** out);

/* UPML state: region index, array 0..7 (vv, vvfo, vvfn, ii, iifo, iifn,
 * voltage flux, current flux), component 0..2.  *n receives the element
 * count, start/size the region box (any may be NULL). */
oe_status oe_ext_upml_read(oe_extension* ext, int region, int which, int component, float* out,
                           size_t capacity, size_t* n, size_t start[3], size_t size[3]);

/* FieldProbe (src/processing/probes.rs:263-303): records Ex..Hz at a node
 * after every step.  Data are read with oe_ext_probe_read. */
oe_status oe_ext_field_probe_create(oe_engine* eng, uint32_t i, uint32_t j, uint32_t k,
                                    oe_extension** out);

/* Number of samples recorded and copy of [n][6] float64 values. */
oe_status oe_ext_probe_read(oe_extension* ext, double* out, size_t capacity, size_t* n_samples);

/* Running DFT of field signals (FrequencyDomainFieldDump::accumulate,
 * src/processing/fields_fd.rs:149-217, with the openEMS pulse convention of
 * python/openEMS/utilities.py:21-35).  Signal m is the float64 sum
 * sum_t weight[t] * F[component[t]][cell[t]] over its terms
 * offsets[m] .. offsets[m+1]; components 0..2 are Ex..Ez, 3..5 Hx..Hz, cell
 * is the linear index (i*ny + j)*nz + k.  A signal's terms must all be E or
 * all H; E signals are sampled at t = (n+1) dt and H signals at
 * (n+1/2) dt after step n.  After every step each signal is accumulated as
 * X(f) += x exp(-j 2 pi f t) (reads return 2 dt X).  precision FP64 keeps
 * complex float64 sums; FP32_KAHAN keeps float32 sums with Kahan
 * compensation.  record_steps > 0 also keeps the raw series. */
typedef enum { OE_DFT_FP64 = 0, OE_DFT_FP32_KAHAN = 1 } oe_dft_precision;
typedef struct {
    size_t n_signals;
    const uint64_t* offsets;   /* n_signals + 1 */
    const uint32_t* component; /* per term */
    const uint64_t* cell;      /* per term */
    const double* weight;      /* per term */
    const double* freqs;       /* Hz */
    size_t n_freqs;
    int precision;             /* oe_dft_precision */
    uint64_t record_steps;
    uint32_t time_stride;      /* accumulate every n-th step (0 or 1: every step); reads scale by n */
} oe_dft_config;
oe_status oe_ext_dft_create(oe_engine* eng, const oe_dft_config* cfg, oe_extension** out);

/* 2 dt X(f) for every (frequency, signal): re/im are [n_freqs][n_signals]. */
oe_status oe_ext_dft_read(oe_extension* ext, double* re, double* im, uint64_t* n_accumulated);

/* Recorded raw series [n][n_signals] (n <= record_steps). */
oe_status oe_ext_dft_read_series(oe_extension* ext, double* out, size_t capacity_steps, size_t* n_steps);

/* Lumped port probes (python/openEMS/ports.py:155-209): a DFT extension with
 * two signals on the port of region/dir (sign = +1 when the port points
 * along +dir): U = -sign * sum(E_dir * d_dir) along the centre line of the
 * cross-section, and I = sign * loop integral of H around the dir-edges of
 * the cross-section at the middle index along dir.  The port's resistor is
 * oe_operator_add_lumped_element; its excitation is a soft source of
 * -sign * amplitude * signal on every dir-edge of region. */
oe_status oe_ext_lumped_port_create(oe_engine* eng, int dir, const size_t region[6], int sign,
                                    const double* freqs, size_t n_freqs, int precision,
                                    uint64_t record_steps, oe_extension** out);

/* Port::CalcPort (python/openEMS/ports.py:115-152) on a lumped port's DFT:
 * uf_inc = (U + I z_ref)/2, uf_ref = U - uf_inc, S11 = uf_ref / uf_inc. */
oe_status oe_ext_port_s11(oe_extension* port, double z_ref, double* s11_re, double* s11_im);

/* ------------------------------------------------------------------ NF2FF
 * The reference's Nf2ff (src/nf2ff/mod.rs:72-170) keeps the API but stubs
 * the surface integral (integrate_surfaces returns zeros, :157-170;
 * max_directivity returns 1, :66-69); the transform here is the one of
 * matlab/AnalyzeNF2FF.m: J = n x H, M = -n x E on every face, radiation
 * integrals with exp(+j k r_hat . r') and 2-D trapezoid weights,
 * E_theta = -jk e^{-jkr}/(4 pi r) (L_phi + Z0 N_theta),
 * E_phi = jk e^{-jkr}/(4 pi r) (L_theta - Z0 N_phi), P_rad = 1/2 sum
 * Re((E x H*) . n) dA, D_max = 4 pi U_max / P_rad.  The per-face contraction
 * over the first in-plane axis runs on the FP32 matrix cores; the phase
 * factors, the second-axis reduction and P_rad are float64. */

/* FarFieldResult (src/nf2ff/mod.rs:11-24) of one frequency.  The field
 * arrays are [n_theta][n_phi], caller-owned, and any may be null: split real
 * and imaginary parts, or e_theta / e_phi as interleaved complex float64
 * ([n_theta][n_phi][re, im], the layout of the reference's Vec<Complex64>;
 * the cheaper form, copied without reordering). */
typedef struct {
    double* e_theta_re;
    double* e_theta_im;
    double* e_phi_re;
    double* e_phi_im;
    double prad;       /* radiated power through the surface, W */
    double dmax;       /* 4 pi U_max / prad (FarFieldResult::max_directivity) */
    double gemm_ms;    /* device time of the matrix-core contractions */
    double total_ms;   /* device time of the whole transform */
    double gemm_flop;  /* real floating-point operations of the contractions */
    double* e_theta;   /* interleaved complex, or null */
    double* e_phi;
    /* The large contractions run on the FP16 matrix cores with each FP32
     * operand split into two fp16 halves (hi + lo, three products per
     * element pair; OE_NF2FF_SPLIT16=0 keeps FP32): their device time and
     * their share of gemm_flop. */
    double gemm_split_ms;
    double gemm_split_flop;
} oe_far_field;

typedef struct oe_nf2ff oe_nf2ff;

/* Nf2ff::new (src/nf2ff/mod.rs:98-105) on HIP device `device`. */
oe_status oe_nf2ff_create(int device, double frequency, oe_nf2ff** out);
void oe_nf2ff_destroy(oe_nf2ff* nf);

/* Nf2ff::add_surface (src/nf2ff/mod.rs:108-110) with a NearFieldSurface
 * (:78-91): the face of normal axis normal_dir (0..2) with outward sign
 * +1/-1 at coordinate `position`, sampled at nodes c1[n1] x c2[n2] along the
 * in-plane axes (normal_dir+1)%3 and (normal_dir+2)%3.  e_t and h_t hold the
 * tangential phasors along those two axes as interleaved complex float64,
 * [2][n1][n2][re, im].  Trapezoid weights follow from c1 and c2. */
oe_status oe_nf2ff_add_surface(oe_nf2ff* nf, int normal_dir, int sign, double position, const double* c1,
                               size_t n1, const double* c2, size_t n2, const double* e_t, const double* h_t);

/* Nf2ff::calculate (src/nf2ff/mod.rs:117-154): far field at distance radius
 * for every (theta[i], phi[j]); center is the phase reference point (null:
 * the origin). */
oe_status oe_nf2ff_calculate(oe_nf2ff* nf, const double* theta, size_t n_theta, const double* phi,
                             size_t n_phi, double radius, const double center[3], oe_far_field* out);

/* NF2FF recording box (openEMS CreateNF2FFBox, python/openEMS/nf2ff.py:26-80,
 * with frequency-domain recording): a running DFT of the tangential E and H
 * phasors on the closed surface through the mesh nodes box = {i0, i1, j0,
 * j1, k0, k1} (inclusive node indices, 1 <= lo < hi <= n-1).  E at a node is
 * the mean of the two edges that meet there, H the mean of the four
 * surrounding faces (openEMS's interpolated dump).  directions is the
 * 6-bit face mask (bit 2*axis + side, side 1 = upper face; 0 = all faces).
 * Node coordinates are index * delta.  time_stride as in oe_dft_config. */
oe_status oe_ext_nf2ff_create(oe_engine* eng, const size_t box[6], int directions, const double* freqs,
                              size_t n_freqs, int precision, uint32_t time_stride, oe_extension** out);

/* Nf2ff::calculate on the recorded phasors of frequency freqs[freq_index]. */
oe_status oe_ext_nf2ff_calculate(oe_extension* ext, size_t freq_index, const double* theta, size_t n_theta,
                                 const double* phi, size_t n_phi, double radius, const double center[3],
                                 oe_far_field* out);

/* ------------------------------------------------------------------ sweep
 * A parameter sweep: many structures stepped together on one device.  No
 * counterpart in the reference, which runs one Simulation at a time
 * (Simulation::run, src/fdtd/simulation.rs:255-326); each structure is an
 * ordinary engine with its CPML (physical mode) and lumped-port DFT, and
 * after a run its fields, port spectra and S11 are read through the usual
 * calls.  Every half-step of all active structures is one launch per kernel
 * variant; the steps between two energy checks are captured as one HIP graph
 * and replayed.  Each structure stops on its own: energy is sampled after
 * every step whose timestep (the engine's count, which continues across
 * runs, as a single engine's does) is a multiple of check_interval (0
 * included); the peak is the maximum of those samples over all of the
 * engine's sweep runs and persistent-peak batches (it is kept in the engine),
 * and the structure stops when the energy falls energy_decay_db below it
 * (the rule of src/fdtd/basic_engine.rs:128-141 with a persistent peak) or
 * after max_steps steps of this run.  So a run of K steps followed by a run
 * to the end gives the same fields, spectra and stopping step as one run.
 * Structures share dt and, at the start of a run, the timestep.  The engines
 * and extensions must outlive the sweep's use of them; the engines must not
 * be used concurrently. */
typedef struct oe_sweep oe_sweep;
typedef struct {
    uint64_t max_steps;
    int has_energy_decay_db;
    double energy_decay_db;
    uint32_t check_interval;  /* 0: 100 */
    int use_graph;            /* replay captured graphs (else plain launches) */
    int autotune;             /* time candidate tile depths on the zero state first */
} oe_sweep_config;
typedef struct {
    uint64_t timesteps_executed;
    int termination_reason;   /* oe_termination_reason */
    double final_decay_db;
    double peak_energy;
    double final_energy;
    double elapsed_seconds;   /* wall time of the whole run */
} oe_sweep_entry;

oe_status oe_sweep_create(int device, oe_sweep** out);
void oe_sweep_destroy(oe_sweep* sw);
/* Adds a structure: its engine, CPML (or null), port DFT (or null; at most
 * 256 signals, every step) and excitations (sample s is applied at step s of
 * each run, counted from the run's start).  *index receives its position. */
oe_status oe_sweep_add(oe_sweep* sw, oe_engine* eng, oe_extension* cpml, oe_extension* port,
                       const oe_excitation* excitations, size_t n_excitations, size_t* index);
/* Runs every structure to its end; results has one entry per structure. */
oe_status oe_sweep_run(oe_sweep* sw, const oe_sweep_config* cfg, oe_sweep_entry* results);
/* The work-group depth along i in use and the autotuner's timings (ms per
 * step of all structures' Yee updates) for its four candidates. */
oe_status oe_sweep_tuning(const oe_sweep* sw, int* ci, double ms_per_step[4], int candidates[4]);
oe_status oe_sweep_set_ci(oe_sweep* sw, int ci);
/* The energies the last run's checks saw for structure `index` (E + H sums
 * of squares, Field3D::energy, src/arrays/field.rs:219-221), with their
 * timesteps, in order: up to capacity entries (either array may be NULL);
 * *n receives the count. */
oe_status oe_sweep_energy_history(const oe_sweep* sw, size_t index, uint64_t* timesteps, double* energies,
                                  size_t capacity, size_t* n);

/* ----------------------------------------------------- learned extrapolator
 * No counterpart in the reference (whose only stop rule is the energy decay
 * test, src/fdtd/basic_engine.rs:128-141).  A 1-D convolution + MLP network
 * maps the early window of a port's V and I (L0 decimated samples, two
 * channels, normalised) to `modes` damped complex exponentials that continue
 * both signals for `tail` decimated samples:
 *   conv (kernel 9, stride 2, padding 4, GELU) 2 -> c1 -> c2 -> c3,
 *   fc1 (GELU) -> hidden, fc2 -> 6 numbers per mode: damping
 *   softplus(o0) * 0.05 per sample, frequency sigmoid(o1) * 0.5 cycles per
 *   sample, V residue o2 + j o3, I residue o4 + j o5;
 *   V(n) = sum_m Re(aV_m exp((-alpha_m + j 2 pi f_m) n)), I likewise.
 * Loss: mean squared error of both tails.  Every matrix product runs on the
 * BF16 matrix cores with FP32 accumulation; training is Adam on FP32 master
 * weights.  Parameters are float32 in the order W1 [c1][2][9], b1, W2
 * [c2][c1][9], b2, W3 [c3][c2][9], b3, W4 [hidden][L3 * c3] (position-major
 * flatten of the last convolution's [L3][c3] output), b4, W5
 * [6 modes][hidden], b5.  Inputs are [batch][L0][2] (V, I interleaved),
 * targets [batch][2][tail]. */
typedef struct oe_extrap oe_extrap;
typedef struct {
    uint32_t L0;       /* window length (decimated samples), multiple of 8 */
    uint32_t c1, c2, c3;
    uint32_t hidden;
    uint32_t modes;    /* <= 256 */
    uint32_t batch;    /* largest batch a call may pass */
    uint32_t tail;     /* predicted samples, <= 4096 */
} oe_extrap_dims;
/* params: n float32 in the order above, or NULL for He-normal weights drawn
 * from seed and zero biases. */
oe_status oe_extrap_create(int device, const oe_extrap_dims* dims, const float* params, uint64_t seed,
                           oe_extrap** out);
void oe_extrap_destroy(oe_extrap* x);
size_t oe_extrap_param_count(const oe_extrap* x);
oe_status oe_extrap_params(oe_extrap* x, float* out, size_t n);
oe_status oe_extrap_set_params(oe_extrap* x, const float* in, size_t n);
/* Modes [batch][6 modes] and (if tails is not NULL) tails [batch][2][tail]. */
oe_status oe_extrap_forward(oe_extrap* x, const float* inputs, size_t batch, float* modes, float* tails);
/* Loss and its gradient for every parameter (same order), no update. */
oe_status oe_extrap_gradients(oe_extrap* x, const float* inputs, const float* targets, size_t batch,
                              double* loss, float* grads);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) at learning rate lr; *loss is the
 * loss before the step. */
oe_status oe_extrap_train_step(oe_extrap* x, const float* inputs, const float* targets, size_t batch,
                               float lr, double* loss);

/* Host-callback extension: the CPU fallback the reference's GPU engine takes
 * for an Extension whose gpu_data() is None (requires_cpu_fallback,
 * src/extensions/extension_trait.rs:143-166), with the hook contract of
 * :32-141.  For oracle and debug use: each selected hook point synchronises
 * the engine's stream, copies the six field arrays (Ex, Ey, Ez, Hx, Hy, Hz,
 * N floats each, z fastest) to a host buffer, calls hook(user, which,
 * timestep, fields) and uploads what that hook may change:
 *   which 0 pre_update_h, 1 post_update_h  -> H uploaded,
 *   which 2 pre_update_e, 3 post_update_e  -> E uploaded,
 *   which 4 pre_batch,    5 post_batch     -> all six uploaded.
 * hook_mask bit `which` selects a hook point.  timestep is the engine's count
 * at the start of the step (current_step, src/fdtd/basic_engine.rs:85-87);
 * for pre/post_batch the count at that moment.  At each hook point the
 * batch's device extensions act first (the order listed at
 * oe_engine_run_batch), then host extensions in list order; sources are
 * applied after the post_update_e hooks.  A nonzero return from hook ends
 * run_batch at once with that status (values outside 1..7 become
 * OE_ERR_NUMERICAL), as the reference's `?` does; post_batch hooks are then
 * not called.  check_termination (may be NULL) is asked at every check step
 * (timestep > 0 and a multiple of check_interval, with or without a decay
 * test); nonzero ends the batch with OE_TERM_EXTENSION_STOP unless the
 * energy decay test ends it at the same step (src/fdtd/basic_engine.rs:
 * 115-147).  It may write a NUL-terminated reason into reason[0 ..
 * reason_capacity), read back with oe_ext_host_callback_reason.  A batch with
 * a host extension launches every step directly (no graphs, no fused steps). */
typedef int (*oe_host_hook_fn)(void* user, int which, uint64_t timestep, float* const fields[6]);
typedef int (*oe_host_check_fn)(void* user, char* reason, size_t reason_capacity);
typedef struct {
    oe_host_hook_fn hook;               /* may be NULL */
    oe_host_check_fn check_termination; /* may be NULL */
    void* user;
    uint32_t hook_mask;
} oe_host_callback_config;
oe_status oe_ext_host_callback_create(oe_engine* eng, const oe_host_callback_config* cfg, oe_extension** out);
/* The reason given by the last check_termination that asked to stop ("" if
 * none); *len (may be NULL) receives its length. */
oe_status oe_ext_host_callback_reason(oe_extension* ext, char* out, size_t capacity, size_t* len);

/* CPML psi arrays / ADE state for inspection: field 0 = E, 1 = H; which 0 =
 * psi_pos, 1 = psi_neg.  Returns the element count in *n. */
oe_status oe_ext_cpml_read_psi(oe_extension* ext, int face, int field, int which, float* out,
                               size_t capacity, size_t* n);
oe_status oe_ext_ade_read_state(oe_extension* ext, float* out, size_t capacity, size_t* n);

/* ------------------------------------------------------- no allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the float64 sums of squares of a caller-owned field
 * block, the computation of oe_engine_energy (Field3D::energy,
 * src/arrays/field.rs:219-221), which allocates nothing.  fields is a device
 * pointer to 6 n floats on `device` (Ex, Ey, Ez, Hx, Hy, Hz, n each: the
 * engine's layout); out is a device pointer to two doubles that receive the E
 * and H sums, written on `stream` (a hipStream_t; NULL is the null stream)
 * without a host synchronisation, so the call can be captured into a graph.
 * Results equal oe_engine_energy's for the same values bit for bit.  The
 * partial sums live in a 16 KB array in the library's device memory, one per
 * GPU, held until the process ends; every call on a GPU shares it, so calls
 * whose launches overlap on one GPU (on different streams, or from several
 * threads) race.  1 <= n <= OE_NOALLOC_MAX_CELLS, otherwise OE_ERR_GRID. */
#define OE_NOALLOC_MAX_CELLS (((uint64_t)1) << 34)
oe_status oe_fields_energy_noalloc(const float* fields, size_t n, double* out, int device, void* stream);

/* Extension::pre_batch/post_batch reset counterpart: zero the state. */
oe_status oe_extension_reset(oe_extension* ext);
void oe_extension_destroy(oe_extension* ext);

#ifdef __cplusplus
}
#endif

#endif /* OE_OE_H */
