// oe_oracle.cpp -- CPU oracle for the FDTD timestep hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
// openems_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
// of bench.py may load it.  No product code path calls it.
//
// It restates the semantics of the Rust reference engine (umgefahren/openems)
// in plain C++, templated on the scalar type:
//   float  -> the FP32 parity oracle (same operation order as BasicEngine,
//             built with -ffp-contract=off, so no FMA contraction exists)
//   double -> the FP64 error oracle (coefficients kept in double).
// The i-slab OpenMP loop mirrors ParallelEngine (src/fdtd/parallel_engine.rs:
// 84-239) and is the timed CPU baseline; each cell is computed independently,
// so the parallel result is bit-identical to the serial one.
//
// Followed reference code (file:line, relative to the reference root):
//   dt, coefficients      src/fdtd/operator.rs:82-141, 187-263
//   H update              src/fdtd/basic_engine.rs:191-254
//   E update              src/fdtd/basic_engine.rs:262-306
//   excitations           src/fdtd/basic_engine.rs:313-395
//   batch loop/energy     src/fdtd/basic_engine.rs:56-163, src/arrays/field.rs:219
//   CPML coefficients     src/extensions/cpml.rs:68-72, 98-177
//   CPML psi recursion    src/extensions/cpml.rs:443-738 (correction disabled
//                         there; "physical" mode adds the Roden-Gedney
//                         correction the reference comments out at :512-513,
//                         :601-602, applied with the freshly updated psi)
//   Mur ABC               src/extensions/mur_abc.rs:88-151, 213-324
//   Lorentz/Drude/Debye   src/extensions/dispersive.rs:134-203, 476-542, 570-637
//                         (compat); "physical" mode uses the textbook ADE
//                         schemes documented at ade_physical() below.
//   field probe           src/processing/probes.rs:289-303
//
// Parity pins: see tests/test_oracle.py (closed forms, known answers from the
// reference's own unit tests, energy invariants, brute-force tiny cases).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <omp.h>
#include <cstddef>
#include <type_traits>

using std::ptrdiff_t;

namespace {

constexpr double C0 = 299792458.0;
constexpr double EPS0 = 8.854187817e-12;
constexpr double MU0 = 1.256637062e-6;
constexpr double Z0 = 376.730313668;

enum { MODE_COMPAT = 0, MODE_PHYSICAL = 1 };
enum { ADE_LORENTZ = 0, ADE_DRUDE = 1, ADE_DEBYE = 2 };

template <class T>
struct CpmlFace {
    bool on = false;
    int dir = 0;
    bool is_min = true;
    size_t L = 0;
    size_t d0 = 0, d1 = 0, d2 = 0;  // psi dims (nx,ny,nz with dir replaced by L)
    std::vector<T> be, bh, ae, ah, kme, kmh;  // kme/kmh = 1/kappa - 1
    std::vector<T> pe_pos, pe_neg, ph_pos, ph_neg;
};

template <class T>
struct MurFace {
    int ny, nyp, nypp;
    bool is_max;
    size_t line, shift;
    size_t s0, s1;
    T coeff;
    std::vector<T> v_nyp, v_nypp;
};

template <class T>
struct Ade {
    int kind = 0, mode = 0;
    std::vector<uint32_t> pos;  // 3 per cell
    size_t n = 0;
    // compat coefficients (reference semantics, f32 for T=float)
    T a = 0, b = 0, c = 0;
    T dt_t = 0;
    // physical coefficients
    double pa = 0, pb = 0, pc = 0;
    std::vector<T> J[3], Jold[3], Eprev[3];
};

// UPML region (src/extensions/pml.rs:131-178): coefficient and flux arrays
// over the region box, z fastest; [3] = component.
template <class T>
struct UpmlRegion {
    size_t start[3], size[3];
    std::vector<T> vv[3], vvfo[3], vvfn[3], ii[3], iifo[3], iifn[3], vflux[3], cflux[3];
    size_t at(size_t i, size_t j, size_t k) const { return (i * size[1] + j) * size[2] + k; }
};

template <class T>
struct Probe {
    uint32_t i, j, k;
    std::vector<double> data;  // 6 per sample
};

// Material operation, kept in order so that a lumped element can find the
// material under each of its edges (the last operation covering the cell).
struct MatOp {
    int pec;
    double eps_r, mu_r, se, sm;
    size_t box[6];
};

// Signals recorded every step: signal m = sum_t w[t] * F[comp[t]][cell[t]]
// (float64 sum in term order), e.g. a port voltage or current, or a field
// sample on an NF2FF surface.
struct SignalSet {
    size_t n;
    std::vector<uint64_t> off;
    std::vector<uint32_t> comp;
    std::vector<uint64_t> cell;
    std::vector<double> w;
    std::vector<double> series;  // [steps][n]
};

// Extension hooks run on the host fields (src/extensions/extension_trait.rs:
// 34-141): hook 0..3 = pre/post_update_h, pre/post_update_e, 4/5 = pre/post
// batch; fields are the six component arrays (T*).  Test infrastructure for
// the host-callback extension of the engine.
typedef int (*oro_hook_fn)(void* user, int hook, uint64_t step, void* const fields[6]);
typedef int (*oro_term_fn)(void* user);

template <class T>
struct Sim {
    size_t nx, ny, nz, N;
    double dx, dy, dz, dt;
    std::vector<T> E[3], H[3], ca[3], cb[3], da[3], db[3];
    uint64_t timestep = 0;
    int cpml_mode = 0;
    bool cpml_on = false;
    CpmlFace<T> cf[6];
    std::vector<MurFace<T>> mur;
    std::vector<UpmlRegion<T>> upml;
    std::vector<Ade<T>> ade;
    std::vector<Probe<T>> probes;
    std::vector<MatOp> ops;
    std::vector<SignalSet> signals;
    int threads = 0;
    oro_hook_fn hook = nullptr;
    oro_term_fn term = nullptr;
    void* hook_user = nullptr;
    unsigned hook_mask = 0;
    int hook_status = 0;  // first nonzero hook return

    size_t at(size_t i, size_t j, size_t k) const { return (i * ny + j) * nz + k; }
};

template <class T>
void init_sim(Sim<T>& s, size_t nx, size_t ny, size_t nz, double dx, double dy, double dz) {
    s.nx = nx; s.ny = ny; s.nz = nz; s.N = nx * ny * nz;
    s.dx = dx; s.dy = dy; s.dz = dz;
    // Operator::calculate_timestep (src/fdtd/operator.rs:82-89)
    double inv_sum = 1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz);
    double dt_max = 1.0 / (C0 * std::sqrt(inv_sum));
    s.dt = 0.99 * dt_max;
    for (int d = 0; d < 3; ++d) {
        s.E[d].assign(s.N, T(0));
        s.H[d].assign(s.N, T(0));
    }
    // init_e/h_coefficients (src/fdtd/operator.rs:92-141)
    double del[3] = {dx, dy, dz};
    for (int d = 0; d < 3; ++d) {
        s.ca[d].assign(s.N, T(1));
        s.da[d].assign(s.N, T(1));
        s.cb[d].assign(s.N, (T)(s.dt / (EPS0 * del[d])));
        s.db[d].assign(s.N, (T)(-(s.dt / (MU0 * del[d]))));
    }
}

// Operator::set_material (src/fdtd/operator.rs:187-244)
template <class T>
void set_material(Sim<T>& s, double eps_r, double mu_r, double sig_e, double sig_m,
                  size_t i0, size_t i1, size_t j0, size_t j1, size_t k0, size_t k1) {
    double dt = s.dt;
    double eps = eps_r * EPS0;
    double loss_e = sig_e * dt / (2.0 * eps);
    T ca_mat = (T)((1.0 - loss_e) / (1.0 + loss_e));
    double cb_factor = (dt / eps) / (1.0 + loss_e);
    T cbv[3] = {(T)(cb_factor / s.dx), (T)(cb_factor / s.dy), (T)(cb_factor / s.dz)};
    double mu = mu_r * MU0;
    double loss_m = sig_m * dt / (2.0 * mu);
    T da_mat = (T)((1.0 - loss_m) / (1.0 + loss_m));
    double db_factor = -(dt / mu) / (1.0 + loss_m);
    T dbv[3] = {(T)(db_factor / s.dx), (T)(db_factor / s.dy), (T)(db_factor / s.dz)};
    s.ops.push_back(MatOp{0, eps_r, mu_r, sig_e, sig_m, {i0, i1, j0, j1, k0, k1}});
    i1 = std::min(i1, s.nx); j1 = std::min(j1, s.ny); k1 = std::min(k1, s.nz);
    for (size_t i = i0; i < i1; ++i)
        for (size_t j = j0; j < j1; ++j)
            for (size_t k = k0; k < k1; ++k) {
                size_t p = s.at(i, j, k);
                for (int d = 0; d < 3; ++d) {
                    s.ca[d][p] = ca_mat; s.cb[d][p] = cbv[d];
                    s.da[d][p] = da_mat; s.db[d][p] = dbv[d];
                }
            }
}

// Operator::set_pec (src/fdtd/operator.rs:249-263)
template <class T>
void set_pec(Sim<T>& s, size_t i0, size_t i1, size_t j0, size_t j1, size_t k0, size_t k1) {
    s.ops.push_back(MatOp{1, 0, 0, 0, 0, {i0, i1, j0, j1, k0, k1}});
    i1 = std::min(i1, s.nx); j1 = std::min(j1, s.ny); k1 = std::min(k1, s.nz);
    for (size_t i = i0; i < i1; ++i)
        for (size_t j = j0; j < j1; ++j)
            for (size_t k = k0; k < k1; ++k) {
                size_t p = s.at(i, j, k);
                for (int d = 0; d < 3; ++d) { s.ca[d][p] = T(0); s.cb[d][p] = T(0); }
            }
}

// Lumped resistor over the dir-edges of a box, as an openEMS LumpedElement
// (python/openEMS/ports.py:177-181): the n_ser edges along dir are in series
// and the n_par edge lines across it in parallel, so each edge carries
// R_edge = R n_par / n_ser, i.e. the conductivity sigma += d_dir / (R_edge A)
// added to the cell's own material before Ca/Cb are formed
// (src/fdtd/operator.rs:199-209).  PEC cells are left PEC.
template <class T>
void add_lumped(Sim<T>& s, int dir, const size_t b[6], double R) {
    double del[3] = {s.dx, s.dy, s.dz};
    size_t n[3] = {b[1] - b[0], b[3] - b[2], b[5] - b[4]};
    size_t n_ser = n[dir], n_par = n[(dir + 1) % 3] * n[(dir + 2) % 3];
    double r_edge = R * (double)n_par / (double)n_ser;
    double area = del[(dir + 1) % 3] * del[(dir + 2) % 3];
    double sig_add = del[dir] / (r_edge * area);
    for (size_t i = b[0]; i < std::min(b[1], s.nx); ++i)
        for (size_t j = b[2]; j < std::min(b[3], s.ny); ++j)
            for (size_t k = b[4]; k < std::min(b[5], s.nz); ++k) {
                double eps_r = 1.0, se = 0.0;
                bool pec = false;
                for (const MatOp& m : s.ops) {
                    if (i >= m.box[0] && i < m.box[1] && j >= m.box[2] && j < m.box[3] && k >= m.box[4] && k < m.box[5]) {
                        if (m.pec) pec = true;
                        else { pec = false; eps_r = m.eps_r; se = m.se; }
                    }
                }
                if (pec) continue;
                double eps = eps_r * EPS0;
                double loss = (se + sig_add) * s.dt / (2.0 * eps);
                size_t p = s.at(i, j, k);
                s.ca[dir][p] = (T)((1.0 - loss) / (1.0 + loss));
                s.cb[dir][p] = (T)((s.dt / eps) / (1.0 + loss) / del[dir]);
            }
}

template <class T>
void record_signals(Sim<T>& s) {
    for (auto& G : s.signals) {
        for (size_t m = 0; m < G.n; ++m) {
            double acc = 0.0;
            for (uint64_t t = G.off[m]; t < G.off[m + 1]; ++t) {
                const uint32_t c = G.comp[t];
                const T* F = c < 3 ? s.E[c].data() : s.H[c - 3].data();
                acc += G.w[t] * (double)F[G.cell[t]];
            }
            G.series.push_back(acc);
        }
    }
}

// ---------------------------------------------------------------- CPML
// CpmlBoundaries::sigma_max + CpmlCoefficients::init (src/extensions/cpml.rs:68-177)
template <class T>
void add_cpml(Sim<T>& s, const size_t layers[6], double m, double R0, double alpha_max,
              double kappa_max, int mode) {
    double del[3] = {s.dx, s.dy, s.dz};
    double dmin = std::min(std::min(s.dx, s.dy), s.dz);
    size_t lmax = 0;
    for (int f = 0; f < 6; ++f) lmax = std::max(lmax, layers[f]);
    double dd = (double)(lmax ? lmax : 1) * dmin;
    double sigma_max = -(m + 1.0) * std::log(R0) / (2.0 * Z0 * dd);
    size_t n[3] = {s.nx, s.ny, s.nz};
    s.cpml_mode = mode;
    for (int f = 0; f < 6; ++f) {
        CpmlFace<T>& F = s.cf[f];
        int dir = f / 2;
        bool is_min = (f % 2) == 0;
        size_t L = layers[f];
        F = CpmlFace<T>();
        if (!(L > 0 && L < n[dir])) continue;
        F.on = true; F.dir = dir; F.is_min = is_min; F.L = L;
        F.be.resize(L); F.bh.resize(L); F.ae.resize(L); F.ah.resize(L);
        F.kme.resize(L); F.kmh.resize(L);
        double delta = del[dir];
        double width = (double)L * delta;
        for (size_t i = 0; i < L; ++i) {
            double depth_e = is_min ? (double)(L - i) * delta : (double)(i + 1) * delta;
            double depth_h = is_min ? (double)(L - i) * delta - 0.5 * delta
                                    : (double)(i + 1) * delta - 0.5 * delta;
            double rho_e = std::min(std::max(depth_e / width, 0.0), 1.0);
            double rho_h = std::min(std::max(depth_h / width, 0.0), 1.0);
            double sig_e = sigma_max * std::pow(rho_e, m);
            double sig_h = sigma_max * std::pow(rho_h, m);
            double al_e = alpha_max * (1.0 - rho_e);
            double al_h = alpha_max * (1.0 - rho_h);
            double ka_e = 1.0 + (kappa_max - 1.0) * std::pow(rho_e, m);
            double ka_h = 1.0 + (kappa_max - 1.0) * std::pow(rho_h, m);
            double b_e = std::exp(-(sig_e / ka_e / EPS0 + al_e / EPS0) * s.dt);
            double b_h = std::exp(-(sig_h / ka_h / EPS0 + al_h / EPS0) * s.dt);
            double sn_e = sig_e / (EPS0 * ka_e), sn_h = sig_h / (EPS0 * ka_h);
            double an_e = al_e / EPS0, an_h = al_h / EPS0;
            double den_e = sn_e + an_e, den_h = sn_h + an_h;
            double a_e = std::fabs(den_e) > 1e-20 ? sn_e / den_e * (1.0 - b_e) : 0.0;
            double a_h = std::fabs(den_h) > 1e-20 ? sn_h / den_h * (1.0 - b_h) : 0.0;
            if (mode == MODE_PHYSICAL) {  // Roden-Gedney a = -a_ref/kappa, sign folded into psi
                a_e /= ka_e;
                a_h /= ka_h;
            }
            F.be[i] = (T)b_e; F.bh[i] = (T)b_h;
            F.ae[i] = (T)a_e; F.ah[i] = (T)a_h;
            F.kme[i] = (T)(1.0 / ka_e - 1.0);
            F.kmh[i] = (T)(1.0 / ka_h - 1.0);
        }
        F.d0 = dir == 0 ? L : s.nx;
        F.d1 = dir == 1 ? L : s.ny;
        F.d2 = dir == 2 ? L : s.nz;
        size_t sz = F.d0 * F.d1 * F.d2;
        F.pe_pos.assign(sz, T(0)); F.pe_neg.assign(sz, T(0));
        F.ph_pos.assign(sz, T(0)); F.ph_neg.assign(sz, T(0));
        s.cpml_on = true;
    }
}

// comp1/comp2 of reference post_update_h/e (src/extensions/cpml.rs:360-364,399-403)
inline void cpml_comps(int dir, int& c1, int& c2) {
    if (dir == 0) { c1 = 1; c2 = 2; }
    else if (dir == 1) { c1 = 2; c2 = 0; }
    else { c1 = 0; c2 = 1; }
}

// Slab pass after the H update (apply_cpml_h_boundary, src/extensions/cpml.rs:443-527;
// differences get_e_differences :621-678).
template <class T>
void cpml_post_h(Sim<T>& s) {
    const size_t nx = s.nx, ny = s.ny, nz = s.nz;
    const T* Ex = s.E[0].data(); const T* Ey = s.E[1].data(); const T* Ez = s.E[2].data();
    for (int dir = 0; dir < 3; ++dir) {
        int c1, c2; cpml_comps(dir, c1, c2);
        for (int side = 0; side < 2; ++side) {
            CpmlFace<T>& F = s.cf[dir * 2 + side];
            if (!F.on) continue;
            size_t L = F.L;
            size_t n_dir = dir == 0 ? nx : (dir == 1 ? ny : nz);
            size_t g0 = F.is_min ? 0 : n_dir - L;
            size_t ilo = 0, ihi = nx, jlo = 0, jhi = ny, klo = 0, khi = nz;
            if (dir == 0) { ilo = g0; ihi = g0 + L; }
            if (dir == 1) { jlo = g0; jhi = g0 + L; }
            if (dir == 2) { klo = g0; khi = g0 + L; }
            T* H1 = s.H[c1].data(); T* H2 = s.H[c2].data();
            const T* D1 = s.db[c1].data(); const T* D2 = s.db[c2].data();
#pragma omp parallel for schedule(static)
            for (ptrdiff_t gi = (ptrdiff_t)ilo; gi < (ptrdiff_t)ihi; ++gi)
                for (size_t gj = jlo; gj < jhi; ++gj)
                    for (size_t gk = klo; gk < khi; ++gk) {
                        size_t p = s.at(gi, gj, gk);
                        size_t l = dir == 0 ? gi - g0 : (dir == 1 ? gj - g0 : gk - g0);
                        size_t li = dir == 0 ? l : gi, lj = dir == 1 ? l : gj, lk = dir == 2 ? l : gk;
                        size_t q = (li * F.d1 + lj) * F.d2 + lk;
                        T dpos, dneg;
                        if (dir == 0) {
                            bool h = (size_t)gi + 1 < nx;
                            dpos = h ? Ez[p + ny * nz] - Ez[p] : T(0);
                            dneg = h ? Ey[p + ny * nz] - Ey[p] : T(0);
                        } else if (dir == 1) {
                            bool h = gj + 1 < ny;
                            dpos = h ? Ex[p + nz] - Ex[p] : T(0);
                            dneg = h ? Ez[p + nz] - Ez[p] : T(0);
                        } else {
                            bool h = gk + 1 < nz;
                            dpos = h ? Ey[p + 1] - Ey[p] : T(0);
                            dneg = h ? Ex[p + 1] - Ex[p] : T(0);
                        }
                        T b = F.bh[l], a = F.ah[l];
                        T pp = b * F.ph_pos[q] + a * dpos;
                        T pn = b * F.ph_neg[q] + a * dneg;
                        F.ph_pos[q] = pp;
                        F.ph_neg[q] = pn;
                        if (s.cpml_mode == MODE_PHYSICAL) {
                            T km = F.kmh[l];
                            H1[p] = H1[p] + D1[p] * (pp - km * dpos);
                            H2[p] = H2[p] + D2[p] * (km * dneg - pn);
                        }
                    }
        }
    }
}

// Slab pass after the E update (apply_cpml_e_boundary, src/extensions/cpml.rs:532-616;
// differences get_h_differences :682-738).  The field correction applies only
// to edges the E update wrote (i, j, k >= 1).
template <class T>
void cpml_post_e(Sim<T>& s) {
    const size_t nx = s.nx, ny = s.ny, nz = s.nz;
    const T* Hx = s.H[0].data(); const T* Hy = s.H[1].data(); const T* Hz = s.H[2].data();
    for (int dir = 0; dir < 3; ++dir) {
        int c1, c2; cpml_comps(dir, c1, c2);
        for (int side = 0; side < 2; ++side) {
            CpmlFace<T>& F = s.cf[dir * 2 + side];
            if (!F.on) continue;
            size_t L = F.L;
            size_t n_dir = dir == 0 ? nx : (dir == 1 ? ny : nz);
            size_t g0 = F.is_min ? 0 : n_dir - L;
            size_t ilo = 0, ihi = nx, jlo = 0, jhi = ny, klo = 0, khi = nz;
            if (dir == 0) { ilo = g0; ihi = g0 + L; }
            if (dir == 1) { jlo = g0; jhi = g0 + L; }
            if (dir == 2) { klo = g0; khi = g0 + L; }
            T* E1 = s.E[c1].data(); T* E2 = s.E[c2].data();
            const T* C1 = s.cb[c1].data(); const T* C2 = s.cb[c2].data();
#pragma omp parallel for schedule(static)
            for (ptrdiff_t gi = (ptrdiff_t)ilo; gi < (ptrdiff_t)ihi; ++gi)
                for (size_t gj = jlo; gj < jhi; ++gj)
                    for (size_t gk = klo; gk < khi; ++gk) {
                        size_t p = s.at(gi, gj, gk);
                        size_t l = dir == 0 ? gi - g0 : (dir == 1 ? gj - g0 : gk - g0);
                        size_t li = dir == 0 ? l : gi, lj = dir == 1 ? l : gj, lk = dir == 2 ? l : gk;
                        size_t q = (li * F.d1 + lj) * F.d2 + lk;
                        T dpos, dneg;
                        if (dir == 0) {
                            bool h = gi > 0;
                            dpos = h ? Hz[p] - Hz[p - ny * nz] : T(0);
                            dneg = h ? Hy[p] - Hy[p - ny * nz] : T(0);
                        } else if (dir == 1) {
                            bool h = gj > 0;
                            dpos = h ? Hx[p] - Hx[p - nz] : T(0);
                            dneg = h ? Hz[p] - Hz[p - nz] : T(0);
                        } else {
                            bool h = gk > 0;
                            dpos = h ? Hy[p] - Hy[p - 1] : T(0);
                            dneg = h ? Hx[p] - Hx[p - 1] : T(0);
                        }
                        T b = F.be[l], a = F.ae[l];
                        T pp = b * F.pe_pos[q] + a * dpos;
                        T pn = b * F.pe_neg[q] + a * dneg;
                        F.pe_pos[q] = pp;
                        F.pe_neg[q] = pn;
                        if (s.cpml_mode == MODE_PHYSICAL && gi > 0 && gj > 0 && gk > 0) {
                            T km = F.kme[l];
                            E1[p] = E1[p] + C1[p] * (pp - km * dpos);
                            E2[p] = E2[p] + C2[p] * (km * dneg - pn);
                        }
                    }
        }
    }
}

// ---------------------------------------------------------------- Yee
// update_h_basic (src/fdtd/basic_engine.rs:191-254).  Forward differences past
// the last index are 0 (the "PEC" of the reference is zero-gradient there).
template <class T>
void update_h(Sim<T>& s) {
    const size_t nx = s.nx, ny = s.ny, nz = s.nz, nyz = ny * nz;
    const T* Ex = s.E[0].data(); const T* Ey = s.E[1].data(); const T* Ez = s.E[2].data();
    T* Hx = s.H[0].data(); T* Hy = s.H[1].data(); T* Hz = s.H[2].data();
    const T *da0 = s.da[0].data(), *da1 = s.da[1].data(), *da2 = s.da[2].data();
    const T *db0 = s.db[0].data(), *db1 = s.db[1].data(), *db2 = s.db[2].data();
#pragma omp parallel for schedule(static)
    for (ptrdiff_t i = 0; i < (ptrdiff_t)nx; ++i) {
        const bool hi = (size_t)i + 1 < nx;
        for (size_t j = 0; j < ny; ++j) {
            const bool hj = j + 1 < ny;
            const size_t base = (size_t)i * nyz + j * nz;
            auto cell = [&](size_t p, bool hk) {
                T dez_dy = hj ? Ez[p + nz] - Ez[p] : T(0);
                T dey_dz = hk ? Ey[p + 1] - Ey[p] : T(0);
                T cx = dez_dy - dey_dz;
                Hx[p] = da0[p] * Hx[p] + db0[p] * cx;
                T dex_dz = hk ? Ex[p + 1] - Ex[p] : T(0);
                T dez_dx = hi ? Ez[p + nyz] - Ez[p] : T(0);
                T cy = dex_dz - dez_dx;
                Hy[p] = da1[p] * Hy[p] + db1[p] * cy;
                T dey_dx = hi ? Ey[p + nyz] - Ey[p] : T(0);
                T dex_dy = hj ? Ex[p + nz] - Ex[p] : T(0);
                T cz = dey_dx - dex_dy;
                Hz[p] = da2[p] * Hz[p] + db2[p] * cz;
            };
            if (nz > 1) {
#pragma GCC ivdep
                for (size_t k = 0; k + 1 < nz; ++k) cell(base + k, true);
            }
            cell(base + nz - 1, false);
        }
    }
}

// update_e_basic (src/fdtd/basic_engine.rs:262-306).  Loops start at 1: the
// index-0 planes of every E component are never written.
template <class T>
void update_e(Sim<T>& s) {
    const size_t nx = s.nx, ny = s.ny, nz = s.nz, nyz = ny * nz;
    T* Ex = s.E[0].data(); T* Ey = s.E[1].data(); T* Ez = s.E[2].data();
    const T* Hx = s.H[0].data(); const T* Hy = s.H[1].data(); const T* Hz = s.H[2].data();
    const T *ca0 = s.ca[0].data(), *ca1 = s.ca[1].data(), *ca2 = s.ca[2].data();
    const T *cb0 = s.cb[0].data(), *cb1 = s.cb[1].data(), *cb2 = s.cb[2].data();
#pragma omp parallel for schedule(static)
    for (ptrdiff_t i = 1; i < (ptrdiff_t)nx; ++i) {
        for (size_t j = 1; j < ny; ++j) {
            const size_t base = (size_t)i * nyz + j * nz;
#pragma GCC ivdep
            for (size_t k = 1; k < nz; ++k) {
                size_t p = base + k;
                T cx = (Hz[p] - Hz[p - nz]) - (Hy[p] - Hy[p - 1]);
                Ex[p] = ca0[p] * Ex[p] + cb0[p] * cx;
                T cy = (Hx[p] - Hx[p - 1]) - (Hz[p] - Hz[p - nyz]);
                Ey[p] = ca1[p] * Ey[p] + cb1[p] * cy;
                T cz = (Hy[p] - Hy[p - nyz]) - (Hx[p] - Hx[p - nz]);
                Ez[p] = ca2[p] * Ez[p] + cb2[p] * cz;
            }
        }
    }
}

// ---------------------------------------------------------------- Mur
// MurFace::new / init_coefficients (src/extensions/mur_abc.rs:88-151)
template <class T>
void add_mur(Sim<T>& s, const int enabled[6], double phase_velocity) {
    size_t n[3] = {s.nx, s.ny, s.nz};
    double del[3] = {s.dx, s.dy, s.dz};
    for (int f = 0; f < 6; ++f) {
        if (!enabled[f]) continue;
        MurFace<T> F;
        F.ny = f / 2; F.is_max = (f % 2) == 1;
        F.nyp = (F.ny + 1) % 3; F.nypp = (F.ny + 2) % 3;
        if (F.is_max) { F.line = n[F.ny] - 1; F.shift = n[F.ny] - 2; }
        else { F.line = 0; F.shift = 1; }
        F.s0 = n[F.nyp]; F.s1 = n[F.nypp];
        double c = phase_velocity > 0 ? phase_velocity : C0;
        double c_dt = c * s.dt;
        F.coeff = (T)((c_dt - del[F.ny]) / (c_dt + del[F.ny]));
        F.v_nyp.assign(F.s0 * F.s1, T(0));
        F.v_nypp.assign(F.s0 * F.s1, T(0));
        s.mur.push_back(F);
    }
}

template <class T>
size_t mur_index(const Sim<T>& s, const MurFace<T>& F, size_t line, size_t a, size_t b) {
    size_t pos[3];
    pos[F.ny] = line; pos[F.nyp] = a; pos[F.nypp] = b;
    return s.at(pos[0], pos[1], pos[2]);
}

// pre_voltage_update (src/extensions/mur_abc.rs:213-254)
template <class T>
void mur_pre(Sim<T>& s) {
    for (auto& F : s.mur) {
        const T* A = s.E[F.nyp].data(); const T* B = s.E[F.nypp].data();
        for (size_t a = 0; a < F.s0; ++a)
            for (size_t b = 0; b < F.s1; ++b) {
                size_t q = a * F.s1 + b;
                size_t pb = mur_index(s, F, F.line, a, b), ps = mur_index(s, F, F.shift, a, b);
                F.v_nyp[q] = A[ps] - F.coeff * A[pb];
                F.v_nypp[q] = B[ps] - F.coeff * B[pb];
            }
    }
}

// post_voltage_update + apply_to_voltages (src/extensions/mur_abc.rs:257-324)
template <class T>
void mur_post_apply(Sim<T>& s) {
    for (auto& F : s.mur) {
        T* A = s.E[F.nyp].data(); T* B = s.E[F.nypp].data();
        for (size_t a = 0; a < F.s0; ++a)
            for (size_t b = 0; b < F.s1; ++b) {
                size_t q = a * F.s1 + b;
                size_t ps = mur_index(s, F, F.shift, a, b);
                F.v_nyp[q] += F.coeff * A[ps];
                F.v_nypp[q] += F.coeff * B[ps];
            }
    }
    for (auto& F : s.mur) {
        T* A = s.E[F.nyp].data(); T* B = s.E[F.nypp].data();
        for (size_t a = 0; a < F.s0; ++a)
            for (size_t b = 0; b < F.s1; ++b) {
                size_t q = a * F.s1 + b;
                size_t pb = mur_index(s, F, F.line, a, b);
                A[pb] = F.v_nyp[q];
                B[pb] = F.v_nypp[q];
            }
    }
}

// ---------------------------------------------------------------- ADE
// params: Lorentz {omega_0, gamma, delta_eps}; Drude {omega_p, gamma, -};
// Debye {tau, delta_eps, -}.  All in SI (rad/s, s).
template <class T>
void add_ade(Sim<T>& s, int kind, const double prm[3], const uint32_t* pos, size_t n, int mode) {
    Ade<T> A;
    A.kind = kind; A.mode = mode; A.n = n;
    A.pos.assign(pos, pos + 3 * n);
    double dt = s.dt;
    if (kind == ADE_LORENTZ) {  // src/extensions/dispersive.rs:134-159
        double w0 = prm[0], g = prm[1], de = prm[2];
        double den = 1.0 + g * dt / 2.0;
        A.a = (T)((2.0 - w0 * w0 * dt * dt) / den);
        A.b = (T)((g * dt / 2.0 - 1.0) / den);
        A.c = (T)(EPS0 * de * w0 * w0 * dt * dt / den);
        A.pa = (2.0 - w0 * w0 * dt * dt) / den;
        A.pb = (g * dt / 2.0 - 1.0) / den;
        A.pc = EPS0 * de * w0 * w0 * dt * dt / den;
    } else if (kind == ADE_DRUDE) {  // :476-498
        double wp = prm[0], g = prm[1];
        double den = 1.0 + g * dt / 2.0;
        A.a = (T)((1.0 - g * dt / 2.0) / den);
        A.b = (T)(EPS0 * wp * wp * dt / den);
        A.c = (T)(dt / EPS0);
        A.pa = (1.0 - g * dt / 2.0) / den;       // kj
        A.pb = (EPS0 * wp * wp * dt / 2.0) / den;  // beta_j
    } else {  // Debye :570-592
        double tau = prm[0], de = prm[1];
        double den = 2.0 * tau + dt;
        A.a = (T)((2.0 * tau - dt) / den);
        A.b = (T)(2.0 * EPS0 * de * dt / den);
        A.c = (T)(dt / EPS0);
        A.pa = (2.0 * tau - dt) / den;
        A.pb = EPS0 * de * dt / den;
    }
    A.dt_t = (T)dt;
    for (int d = 0; d < 3; ++d) {
        A.J[d].assign(n, T(0)); A.Jold[d].assign(n, T(0)); A.Eprev[d].assign(n, T(0));
    }
    s.ade.push_back(std::move(A));
}

template <class T>
void ade_pre(Sim<T>& s) {
    for (auto& A : s.ade) {
        if (A.mode != MODE_PHYSICAL) continue;
        for (size_t c = 0; c < A.n; ++c) {
            size_t p = s.at(A.pos[3 * c], A.pos[3 * c + 1], A.pos[3 * c + 2]);
            for (int d = 0; d < 3; ++d) A.Eprev[d][c] = s.E[d][p];
        }
    }
}

// Compat: the reference's post_update_e bodies verbatim in arithmetic.
// Physical (per cell and component, K = Cb*delta = dt/eps_cell/(1+loss)):
//   Drude, trapezoidal J-E coupling (unconditionally stable):
//     J^{n+1} = kj J^n + bj (E^{n+1} + E^n)
//     E^{n+1} = [E* - K/2 ((1+kj) J^n + bj E^n)] / (1 + K bj / 2)
//   Lorentz, second-order P recursion (keeps the P^{n-1} term):
//     P^{n+1} = a P^n + b P^{n-1} + c E^n ;  E^{n+1} = E* - K (P^{n+1}-P^n)/dt
//   Debye, trapezoidal:  P^{n+1} = a P^n + b (E^{n+1} + E^n)
//     E^{n+1} = [E* - K/dt ((a-1) P^n + b E^n)] / (1 + K b / dt)
// E* is the field after the Yee update, E^n the field captured before it.
template <class T>
void ade_post(Sim<T>& s) {
    double del[3] = {s.dx, s.dy, s.dz};
    for (auto& A : s.ade) {
        for (size_t c = 0; c < A.n; ++c) {
            size_t p = s.at(A.pos[3 * c], A.pos[3 * c + 1], A.pos[3 * c + 2]);
            for (int d = 0; d < 3; ++d) {
                T e = s.E[d][p];
                if (A.mode == MODE_COMPAT) {
                    T j_old = A.J[d][c];
                    if (A.kind == ADE_LORENTZ) {
                        T jn = A.a * j_old + A.c * e;
                        A.J[d][c] = jn;
                        s.E[d][p] = e - A.b * jn;
                    } else if (A.kind == ADE_DRUDE) {
                        T jn = A.a * j_old + A.b * e;
                        A.J[d][c] = jn;
                        s.E[d][p] = e - A.c * jn;
                    } else {
                        T pn = A.a * j_old + A.b * e;
                        A.J[d][c] = pn;
                        T dpdt = (pn - j_old) / A.dt_t;
                        s.E[d][p] = e - A.c * dpdt;
                    }
                } else {
                    T K = (T)((double)s.cb[d][p] * del[d]);
                    T en = A.Eprev[d][c];
                    T jn = A.J[d][c];
                    if (A.kind == ADE_DRUDE) {
                        T kj = (T)A.pa, bj = (T)A.pb;
                        T enew = (e - (K / T(2)) * ((T(1) + kj) * jn + bj * en)) /
                                 (T(1) + K * bj / T(2));
                        A.J[d][c] = kj * jn + bj * (enew + en);
                        s.E[d][p] = enew;
                    } else if (A.kind == ADE_LORENTZ) {
                        T pold = A.Jold[d][c];
                        T pnew = (T)A.pa * jn + (T)A.pb * pold + (T)A.pc * en;
                        A.Jold[d][c] = jn;
                        A.J[d][c] = pnew;
                        s.E[d][p] = e - K * ((pnew - jn) / (T)s.dt);
                    } else {
                        T a = (T)A.pa, b = (T)A.pb, dtt = (T)s.dt;
                        T enew = (e - (K / dtt) * ((a - T(1)) * jn + b * en)) /
                                 (T(1) + K * b / dtt);
                        A.J[d][c] = a * jn + b * (enew + en);
                        s.E[d][p] = enew;
                    }
                }
            }
        }
    }
}

template <class T>
double energy_of(const Sim<T>& s, const std::vector<T>* F) {
    double tot = 0.0;
    for (int d = 0; d < 3; ++d) {
        const T* v = F[d].data();
        double acc = 0.0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
        for (ptrdiff_t p = 0; p < (ptrdiff_t)s.N; ++p) acc += (double)v[p] * (double)v[p];
        tot += acc;
    }
    return tot;
}

// Excitation table entry (ScheduledExcitation, src/fdtd/batch.rs:71-83)
// ---------------------------------------------------------------- UPML
// Upml::calculate_kappa (src/extensions/pml.rs:408-461): grading from the
// outer boundary of the region's own direction; the current value is taken
// half a cell inside (component == direction) or outside.
inline void upml_kappa(size_t n_dir, int component, int direction, size_t pos, bool is_min, size_t layers,
                       double delta, double width, double sigma_max, double m, double* kv, double* ki) {
    double depth = 0.0;
    if (is_min) {
        if (pos <= layers) depth = (double)(layers - pos) * delta;
    } else {
        const size_t bs = n_dir - layers - 1;
        if (pos >= bs) depth = (double)(pos - bs) * delta;
    }
    if (depth <= 0.0) { *kv = 0.0; *ki = 0.0; return; }
    *kv = sigma_max * std::pow(depth / width, m);
    const double di = depth + (component == direction ? -0.5 : 0.5) * delta;
    double nd = di / width;
    nd = nd < 0.0 ? 0.0 : (nd > 1.0 ? 1.0 : nd);
    *ki = sigma_max * std::pow(nd, m);
}

// Upml::build_regions + init_region_coefficients (src/extensions/pml.rs:222-404)
template <class T>
void add_upml(Sim<T>& s, const size_t layers[6], double m, double R) {
    const size_t n[3] = {s.nx, s.ny, s.nz};
    const double del[3] = {s.dx, s.dy, s.dz};
    auto region = [&](size_t st0, size_t st1, size_t st2, size_t sz0, size_t sz1, size_t sz2, int dir, bool is_min) {
        UpmlRegion<T> r;
        r.start[0] = st0; r.start[1] = st1; r.start[2] = st2;
        r.size[0] = sz0; r.size[1] = sz1; r.size[2] = sz2;
        const size_t cnt = sz0 * sz1 * sz2;
        for (int c = 0; c < 3; ++c) {
            for (auto* v : {&r.vv[c], &r.vvfo[c], &r.vvfn[c], &r.ii[c], &r.iifo[c], &r.iifn[c], &r.vflux[c], &r.cflux[c]})
                v->assign(cnt, T(0));
        }
        const size_t L = is_min ? layers[2 * dir] : layers[2 * dir + 1];
        const double delta = del[dir], width = (double)L * delta;
        const double smax = -(m + 1.0) * std::log(R) / (2.0 * Z0 * width);
        for (size_t li = 0; li < sz0; ++li)
            for (size_t lj = 0; lj < sz1; ++lj)
                for (size_t lk = 0; lk < sz2; ++lk) {
                    const size_t pos[3] = {li + st0, lj + st1, lk + st2};
                    double kv[3], ki[3];
                    for (int c = 0; c < 3; ++c)
                        upml_kappa(n[dir], c, dir, pos[dir], is_min, L, delta, width, smax, m, &kv[c], &ki[c]);
                    const size_t q = r.at(li, lj, lk);
                    for (int c = 0; c < 3; ++c) {
                        const int npp = (c + 2) % 3;
                        if (kv[0] + kv[1] + kv[2] > 0.0) {  // eq. 7.88
                            const double den = 2.0 * EPS0 + kv[npp] * s.dt;
                            r.vv[c][q] = (T)((2.0 * EPS0 - kv[npp] * s.dt) / den);
                            r.vvfn[c][q] = (T)((2.0 * EPS0 + kv[c] * s.dt) / den);
                            r.vvfo[c][q] = (T)((2.0 * EPS0 - kv[c] * s.dt) / den);
                        } else {
                            r.vv[c][q] = T(1); r.vvfn[c][q] = T(1); r.vvfo[c][q] = T(0);
                        }
                        if (ki[0] + ki[1] + ki[2] > 0.0) {  // eq. 7.90 (with eps0, as the reference)
                            const double den = 2.0 * EPS0 + ki[npp] * s.dt;
                            r.ii[c][q] = (T)((2.0 * EPS0 - ki[npp] * s.dt) / den);
                            r.iifn[c][q] = (T)((2.0 * EPS0 + ki[c] * s.dt) / den);
                            r.iifo[c][q] = (T)((2.0 * EPS0 - ki[c] * s.dt) / den);
                        } else {
                            r.ii[c][q] = T(1); r.iifn[c][q] = T(1); r.iifo[c][q] = T(0);
                        }
                    }
                }
        s.upml.push_back(std::move(r));
    };
    const size_t nx = n[0], ny = n[1], nz = n[2];
    if (layers[0] > 0 && layers[0] < nx) region(0, 0, 0, layers[0] + 1, ny, nz, 0, true);
    if (layers[1] > 0 && layers[1] < nx) region(nx - layers[1] - 1, 0, 0, layers[1] + 1, ny, nz, 0, false);
    const size_t xs = layers[0] > 0 ? layers[0] + 1 : 0, xe = layers[1] > 0 ? nx - layers[1] - 1 : nx;
    if (layers[2] > 0 && layers[2] < ny && xs < xe) region(xs, 0, 0, xe - xs, layers[2] + 1, nz, 1, true);
    if (layers[3] > 0 && layers[3] < ny && xs < xe) region(xs, ny - layers[3] - 1, 0, xe - xs, layers[3] + 1, nz, 1, false);
    const size_t ys = layers[2] > 0 ? layers[2] + 1 : 0, ye = layers[3] > 0 ? ny - layers[3] - 1 : ny;
    if (layers[4] > 0 && layers[4] < nz && xs < xe && ys < ye)
        region(xs, ys, 0, xe - xs, ye - ys, layers[4] + 1, 2, true);
    if (layers[5] > 0 && layers[5] < nz && xs < xe && ys < ye)
        region(xs, ys, nz - layers[5] - 1, xe - xs, ye - ys, layers[5] + 1, 2, false);
}

// pre_voltage_update / pre_current_update (src/extensions/pml.rs:500-537, 584-619):
// f_help = c * F - cfo * flux; F <- flux; flux <- f_help
template <class T>
void upml_pre(Sim<T>& s, bool e) {
    for (auto& r : s.upml)
        for (size_t li = 0; li < r.size[0]; ++li)
            for (size_t lj = 0; lj < r.size[1]; ++lj)
                for (size_t lk = 0; lk < r.size[2]; ++lk) {
                    const size_t q = r.at(li, lj, lk);
                    const size_t p = s.at(li + r.start[0], lj + r.start[1], lk + r.start[2]);
                    for (int c = 0; c < 3; ++c) {
                        T* F = e ? s.E[c].data() : s.H[c].data();
                        std::vector<T>& fl = e ? r.vflux[c] : r.cflux[c];
                        const T cc = e ? r.vv[c][q] : r.ii[c][q];
                        const T cfo = e ? r.vvfo[c][q] : r.iifo[c][q];
                        const T old = F[p], fo = fl[q];
                        const T fh = cc * old - cfo * fo;
                        F[p] = fo;
                        fl[q] = fh;
                    }
                }
}

// post_voltage_update / post_current_update (src/extensions/pml.rs:540-580, 622-657):
// F_new = f_help + cfn * flux_new; flux <- flux_new (the engine's update of F)
template <class T>
void upml_post(Sim<T>& s, bool e) {
    for (auto& r : s.upml)
        for (size_t li = 0; li < r.size[0]; ++li)
            for (size_t lj = 0; lj < r.size[1]; ++lj)
                for (size_t lk = 0; lk < r.size[2]; ++lk) {
                    const size_t q = r.at(li, lj, lk);
                    const size_t p = s.at(li + r.start[0], lj + r.start[1], lk + r.start[2]);
                    for (int c = 0; c < 3; ++c) {
                        T* F = e ? s.E[c].data() : s.H[c].data();
                        std::vector<T>& fl = e ? r.vflux[c] : r.cflux[c];
                        const T cfn = e ? r.vvfn[c][q] : r.iifn[c][q];
                        const T fh = fl[q], fn = F[p];
                        fl[q] = fn;
                        F[p] = fh + cfn * fn;
                    }
                }
}

struct Exc {
    uint32_t i, j, k, dir, soft;
};

// run_batch (src/fdtd/basic_engine.rs:56-163) with the hook order
//   pre_h(UPML), H, post_h(UPML, CPML) | pre_e(Mur, ADE, UPML), E,
//   post_e(UPML, CPML, ADE, Mur) | excitations
// A host hook runs after the restated extensions of its hook point, as an
// extension listed after them would.
template <class T>
void call_hook(Sim<T>& s, int which) {
    if (!s.hook || !(s.hook_mask & (1u << which)) || s.hook_status) return;
    void* f[6] = {s.E[0].data(), s.E[1].data(), s.E[2].data(), s.H[0].data(), s.H[1].data(), s.H[2].data()};
    s.hook_status = s.hook(s.hook_user, which, s.timestep, f);
}

template <class T>
void step(Sim<T>& s) {
    if (!s.upml.empty()) upml_pre(s, false);
    call_hook(s, 0);
    update_h(s);
    if (!s.upml.empty()) upml_post(s, false);
    if (s.cpml_on) cpml_post_h(s);
    call_hook(s, 1);
    mur_pre(s);
    ade_pre(s);
    if (!s.upml.empty()) upml_pre(s, true);
    call_hook(s, 2);
    update_e(s);
    if (!s.upml.empty()) upml_post(s, true);
    if (s.cpml_on) cpml_post_e(s);
    ade_post(s);
    mur_post_apply(s);
    call_hook(s, 3);
}

template <class T>
void apply_exc(Sim<T>& s, const Exc* ex, size_t n_exc, const double* samples, size_t n_samples,
               uint64_t step_in_batch) {
    for (size_t e = 0; e < n_exc; ++e) {
        if (step_in_batch >= n_samples) continue;  // Sampled(...).get(step) == None
        T v = (T)samples[e * n_samples + step_in_batch];
        size_t p = s.at(ex[e].i, ex[e].j, ex[e].k);
        T* F = s.E[ex[e].dir].data();
        if (ex[e].soft) F[p] = F[p] + v;
        else F[p] = v;
    }
}

template <class T>
void sample_probes(Sim<T>& s) {
    for (auto& P : s.probes) {
        size_t p = s.at(P.i, P.j, P.k);
        for (int d = 0; d < 3; ++d) P.data.push_back((double)s.E[d][p]);
        for (int d = 0; d < 3; ++d) P.data.push_back((double)s.H[d][p]);
    }
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

struct oro_batch {
    uint64_t num_steps;            // 0 => use max_timesteps (Option<u64>::None)
    uint64_t max_timesteps;        // 0 => None
    const Exc* exc;
    size_t n_exc;
    const double* samples;         // [n_exc][n_samples]
    size_t n_samples;
    uint64_t check_interval;       // default 100
    double energy_decay_db;        // <=0 => None
    uint64_t sample_interval;      // 0 => disabled
    int track_peak;
    // outputs
    double* energy_out;            // [cap][3]: timestep, e_energy, h_energy
    size_t energy_cap;
    size_t n_energy;
    uint64_t steps_executed;
    int reason;                    // 0 steps, 1 energy decay, 2 extension stop
    double final_decay_db;
};

struct oro_handle {
    int prec;
    Sim<float>* f;
    Sim<double>* d;
};

#define DISPATCH(h, ...)               \
    do {                               \
        if ((h)->prec == 32) {         \
            auto& s = *(h)->f;         \
            __VA_ARGS__;               \
        } else {                       \
            auto& s = *(h)->d;         \
            __VA_ARGS__;               \
        }                              \
    } while (0)

void* oro_create(int prec, size_t nx, size_t ny, size_t nz, double dx, double dy, double dz) {
    oro_handle* h = new oro_handle();
    h->prec = prec;
    h->f = nullptr; h->d = nullptr;
    if (prec == 32) { h->f = new Sim<float>(); init_sim(*h->f, nx, ny, nz, dx, dy, dz); }
    else { h->d = new Sim<double>(); init_sim(*h->d, nx, ny, nz, dx, dy, dz); }
    return h;
}

void oro_destroy(void* hv) {
    oro_handle* h = (oro_handle*)hv;
    delete h->f; delete h->d; delete h;
}

void oro_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int oro_max_threads(void) { return omp_get_max_threads(); }

double oro_dt(void* hv) { oro_handle* h = (oro_handle*)hv; double r = 0; DISPATCH(h, r = s.dt); return r; }
uint64_t oro_timestep(void* hv) { oro_handle* h = (oro_handle*)hv; uint64_t r = 0; DISPATCH(h, r = s.timestep); return r; }

void oro_set_material(void* hv, double eps_r, double mu_r, double se, double sm, size_t i0, size_t i1,
                      size_t j0, size_t j1, size_t k0, size_t k1) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, set_material(s, eps_r, mu_r, se, sm, i0, i1, j0, j1, k0, k1));
}

void oro_set_pec(void* hv, size_t i0, size_t i1, size_t j0, size_t j1, size_t k0, size_t k1) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, set_pec(s, i0, i1, j0, j1, k0, k1));
}

// which: 0..2 ca, 3..5 cb, 6..8 da, 9..11 db
void oro_get_coeff(void* hv, int which, double* out) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, {
        auto& v = which < 3 ? s.ca[which] : which < 6 ? s.cb[which - 3] : which < 9 ? s.da[which - 6] : s.db[which - 9];
        for (size_t p = 0; p < s.N; ++p) out[p] = (double)v[p];
    });
}

void oro_get_coeff_f32(void* hv, int which, float* out) {
    oro_handle* h = (oro_handle*)hv;
    if (h->prec != 32) return;
    auto& s = *h->f;
    auto& v = which < 3 ? s.ca[which] : which < 6 ? s.cb[which - 3] : which < 9 ? s.da[which - 6] : s.db[which - 9];
    std::memcpy(out, v.data(), s.N * sizeof(float));
}

// Coefficient arrays handed over whole, as EngineImpl::new clones them from
// the Operator (src/fdtd/basic_engine.rs:42-54); FP32 oracle only.
void oro_set_coeff_f32(void* hv, int which, const float* in) {
    oro_handle* h = (oro_handle*)hv;
    if (h->prec != 32) return;
    auto& s = *h->f;
    auto& v = which < 3 ? s.ca[which] : which < 6 ? s.cb[which - 3] : which < 9 ? s.da[which - 6] : s.db[which - 9];
    std::memcpy(v.data(), in, s.N * sizeof(float));
}

// fields: [Ex,Ey,Ez,Hx,Hy,Hz] each N
void oro_get_fields(void* hv, double* out) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, {
        for (int d = 0; d < 3; ++d)
            for (size_t p = 0; p < s.N; ++p) { out[d * s.N + p] = s.E[d][p]; out[(3 + d) * s.N + p] = s.H[d][p]; }
    });
}

void oro_get_fields_f32(void* hv, float* out) {
    oro_handle* h = (oro_handle*)hv;
    if (h->prec != 32) return;
    auto& s = *h->f;
    for (int d = 0; d < 3; ++d) {
        std::memcpy(out + d * s.N, s.E[d].data(), s.N * sizeof(float));
        std::memcpy(out + (3 + d) * s.N, s.H[d].data(), s.N * sizeof(float));
    }
}

void oro_set_fields(void* hv, const double* in) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, {
        using TT = typename std::remove_reference<decltype(s.E[0][0])>::type;
        for (int d = 0; d < 3; ++d)
            for (size_t p = 0; p < s.N; ++p) { s.E[d][p] = (TT)in[d * s.N + p]; s.H[d][p] = (TT)in[(3 + d) * s.N + p]; }
    });
}

void oro_reset(void* hv) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, {
        for (int d = 0; d < 3; ++d) { std::fill(s.E[d].begin(), s.E[d].end(), 0); std::fill(s.H[d].begin(), s.H[d].end(), 0); }
        s.timestep = 0;
    });
}

void oro_add_cpml(void* hv, const size_t layers[6], double m, double R0, double alpha_max, double kappa_max, int mode) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, add_cpml(s, layers, m, R0, alpha_max, kappa_max, mode));
}

// returns psi array size; field 0 = E, 1 = H; which 0 = pos, 1 = neg; 2..5 -> coefficient tables b,a,km (len L)
size_t oro_cpml_psi(void* hv, int face, int field, int which, double* out) {
    oro_handle* h = (oro_handle*)hv;
    size_t n = 0;
    DISPATCH(h, {
        auto& F = s.cf[face];
        if (F.on) {
            const auto* v = field == 0 ? (which == 0 ? &F.pe_pos : which == 1 ? &F.pe_neg : which == 2 ? &F.be : which == 3 ? &F.ae : &F.kme)
                                       : (which == 0 ? &F.ph_pos : which == 1 ? &F.ph_neg : which == 2 ? &F.bh : which == 3 ? &F.ah : &F.kmh);
            n = v->size();
            if (out) for (size_t q = 0; q < n; ++q) out[q] = (double)(*v)[q];
        }
    });
    return n;
}

void oro_add_upml(void* hv, const size_t layers[6], double m, double R) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, add_upml(s, layers, m, R));
}

// UPML region q: which 0..5 = vv, vvfo, vvfn, ii, iifo, iifn; 6 = volt flux,
// 7 = current flux; component c.  Returns the element count; start/size
// receive the region box.
size_t oro_upml_array(void* hv, size_t q, int which, int c, double* out, size_t* start, size_t* size) {
    oro_handle* h = (oro_handle*)hv;
    size_t n = 0;
    DISPATCH(h, {
        if (q < s.upml.size()) {
            auto& r = s.upml[q];
            const auto* arrs = which == 0 ? &r.vv[c] : which == 1 ? &r.vvfo[c] : which == 2 ? &r.vvfn[c]
                             : which == 3 ? &r.ii[c] : which == 4 ? &r.iifo[c] : which == 5 ? &r.iifn[c]
                             : which == 6 ? &r.vflux[c] : &r.cflux[c];
            n = arrs->size();
            if (out) for (size_t k = 0; k < n; ++k) out[k] = (double)(*arrs)[k];
            if (start) for (int d = 0; d < 3; ++d) start[d] = r.start[d];
            if (size) for (int d = 0; d < 3; ++d) size[d] = r.size[d];
        }
    });
    return n;
}

size_t oro_upml_regions(void* hv) {
    oro_handle* h = (oro_handle*)hv;
    size_t n = 0;
    DISPATCH(h, n = s.upml.size());
    return n;
}

void oro_add_mur(void* hv, const int enabled[6], double phase_velocity) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, add_mur(s, enabled, phase_velocity));
}

void oro_add_ade(void* hv, int kind, const double prm[3], const uint32_t* pos, size_t n, int mode) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, add_ade(s, kind, prm, pos, n, mode));
}

// state of ADE object idx: out [3][n] of J (or P)
size_t oro_ade_state(void* hv, size_t idx, double* out) {
    oro_handle* h = (oro_handle*)hv;
    size_t n = 0;
    DISPATCH(h, {
        if (idx < s.ade.size()) {
            auto& A = s.ade[idx];
            n = A.n;
            if (out) for (int d = 0; d < 3; ++d) for (size_t c = 0; c < n; ++c) out[d * n + c] = (double)A.J[d][c];
        }
    });
    return n;
}

size_t oro_add_field_probe(void* hv, uint32_t i, uint32_t j, uint32_t k) {
    oro_handle* h = (oro_handle*)hv;
    size_t idx = 0;
    DISPATCH(h, { Probe<typename std::remove_reference<decltype(s.E[0][0])>::type> P; P.i = i; P.j = j; P.k = k; s.probes.push_back(P); idx = s.probes.size() - 1; });
    return idx;
}

// returns number of samples; out gets [n][6]
size_t oro_probe_data(void* hv, size_t idx, double* out) {
    oro_handle* h = (oro_handle*)hv;
    size_t n = 0;
    DISPATCH(h, {
        if (idx < s.probes.size()) {
            auto& P = s.probes[idx];
            n = P.data.size() / 6;
            if (out) std::memcpy(out, P.data.data(), P.data.size() * sizeof(double));
        }
    });
    return n;
}

void oro_add_lumped(void* hv, int dir, const size_t box[6], double R) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, add_lumped(s, dir, box, R));
}

size_t oro_add_signals(void* hv, size_t n, const uint64_t* off, const uint32_t* comp, const uint64_t* cell,
                       const double* w) {
    oro_handle* h = (oro_handle*)hv;
    size_t idx = 0;
    DISPATCH(h, {
        SignalSet G;
        G.n = n;
        G.off.assign(off, off + n + 1);
        size_t nt = off[n];
        G.comp.assign(comp, comp + nt);
        G.cell.assign(cell, cell + nt);
        G.w.assign(w, w + nt);
        s.signals.push_back(std::move(G));
        idx = s.signals.size() - 1;
    });
    return idx;
}

// number of recorded steps; out gets [steps][n]
size_t oro_signal_series(void* hv, size_t idx, double* out) {
    oro_handle* h = (oro_handle*)hv;
    size_t steps = 0;
    DISPATCH(h, {
        if (idx < s.signals.size()) {
            auto& G = s.signals[idx];
            steps = G.n ? G.series.size() / G.n : 0;
            if (out) std::memcpy(out, G.series.data(), G.series.size() * sizeof(double));
        }
    });
    return steps;
}

double oro_energy(void* hv, int which) {
    oro_handle* h = (oro_handle*)hv;
    double r = 0;
    DISPATCH(h, r = which == 0 ? energy_of(s, s.E) : energy_of(s, s.H));
    return r;
}

// Run only the Yee H+E updates (no hooks, no excitation): the CPU-baseline kernel.
void oro_yee_steps(void* hv, uint64_t n) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, {
        for (uint64_t t = 0; t < n; ++t) step(s);
    });
}

// EngineImpl::run_batch restated (src/fdtd/basic_engine.rs:56-163)
int oro_run_batch(void* hv, oro_batch* b) {
    oro_handle* h = (oro_handle*)hv;
    int rc = 0;
    DISPATCH(h, {
        uint64_t num_steps = b->num_steps ? b->num_steps : (b->max_timesteps ? b->max_timesteps : UINT64_MAX);
        double peak = 0.0;
        b->n_energy = 0; b->steps_executed = 0; b->reason = 0; b->final_decay_db = 0.0;
        uint64_t ci = b->check_interval ? b->check_interval : 100;
        s.hook_status = 0;
        call_hook(s, 4);
        for (uint64_t st = 0; st < num_steps && !s.hook_status; ++st) {
            uint64_t cur = s.timestep;
            step(s);
            if (s.hook_status) break;
            apply_exc(s, b->exc, b->n_exc, b->samples, b->n_samples, st);
            sample_probes(s);
            record_signals(s);
            s.timestep += 1;
            b->steps_executed += 1;
            if (b->sample_interval > 0 && cur % b->sample_interval == 0) {
                double ee = energy_of(s, s.E), he = energy_of(s, s.H);
                if (b->track_peak) peak = std::max(peak, ee + he);
                if (b->n_energy < b->energy_cap && b->energy_out) {
                    b->energy_out[3 * b->n_energy] = (double)cur;
                    b->energy_out[3 * b->n_energy + 1] = ee;
                    b->energy_out[3 * b->n_energy + 2] = he;
                }
                b->n_energy++;
            }
            if (cur > 0 && cur % ci == 0) {
                // check_termination first, the decay test overrides it (basic_engine.rs:115-147)
                const bool ext_stop = s.term && s.term(s.hook_user);
                if (b->energy_decay_db > 0) {
                    double curE = energy_of(s, s.E) + energy_of(s, s.H);
                    if (peak > 0.0 && curE > 0.0) {
                        double ddb = 10.0 * std::log10(curE / peak);
                        if (ddb < -std::fabs(b->energy_decay_db)) {
                            b->reason = 1; b->final_decay_db = ddb;
                            break;
                        }
                    }
                }
                if (ext_stop) {
                    b->reason = 2;
                    break;
                }
            }
        }
        call_hook(s, 5);
        rc = s.hook_status;
    });
    return rc;
}

void oro_set_hook(void* hv, oro_hook_fn fn, oro_term_fn term, void* user, unsigned mask) {
    oro_handle* h = (oro_handle*)hv;
    DISPATCH(h, {
        s.hook = fn;
        s.term = term;
        s.hook_user = user;
        s.hook_mask = mask;
    });
}

}  // extern "C"
