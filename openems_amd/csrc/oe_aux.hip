// oe_aux.hip -- small gfx950 kernels around the Yee updates: coefficient
// fills, excitations + probes, the float64 energy reduction, Mur ABC faces
// and the dispersive (ADE) polarisation-current updates.

#include "oe_internal.h"

#pragma clang fp contract(off)

namespace oe {
namespace {

__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Operator::set_material / set_pec region writes (src/fdtd/operator.rs:222-262),
// many in one launch: boxes that are pairwise disjoint, each
// writing the coefficient arrays of its mask (bit a: array a of the 12).
// Work-group w fills cells [(w - wg_first[o]) BF_CELLS, ...) of box o, the
// last box whose wg_first is at most w.
constexpr int BF_CELLS = 2048;
__global__ __launch_bounds__(256) void k_boxes_fill(float* __restrict__ Cf, size_t N, int ny, int nz,
                                                    const BoxFill* __restrict__ ops,
                                                    const int* __restrict__ wg_first, int n_ops) {
    int lo = 0, hi = n_ops - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (wg_first[mid] <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const BoxFill& o = ops[lo];
    const int ni = o.box[1] - o.box[0], nj = o.box[3] - o.box[2], nk = o.box[5] - o.box[4];
    const size_t cnt = (size_t)ni * nj * nk;
    const size_t c0 = (size_t)(blockIdx.x - wg_first[lo]) * BF_CELLS;
    const size_t c1 = c0 + BF_CELLS < cnt ? c0 + BF_CELLS : cnt;
    for (size_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const int k = (int)(c % nk);
        const size_t r = c / nk;
        const int j = (int)(r % nj);
        const int i = (int)(r / nj);
        const size_t p = ((size_t)(o.box[0] + i) * ny + (o.box[2] + j)) * nz + (o.box[4] + k);
        for (int a = 0; a < 12; ++a)
            if ((o.mask >> a) & 1u) Cf[(size_t)a * N + p] = o.v[a];
    }
}

// Excitations (BasicEngine::apply_excitations, src/fdtd/basic_engine.rs:313-395)
// run in list order on one lane, as repeated sources on one edge must add up
// in that order; then each probe lane records its node (FieldProbe::sample,
// src/processing/probes.rs:289-303).
__global__ void k_point_ops(float* __restrict__ F, size_t N, const PointExc* __restrict__ exc, int n_exc,
                            uint64_t step, const PointProbe* __restrict__ probes, int n_probes,
                            const uint64_t* __restrict__ step_dev) {
    if (step_dev) step = *step_dev;
    if (threadIdx.x == 0) {
        for (int e = 0; e < n_exc; ++e) {
            const PointExc x = exc[e];
            if (step >= x.n_samples) continue;
            const float v = x.samples[step];
            float* f = F + (size_t)x.comp * N + x.idx;
            if (x.soft) *f = *f + v;
            else *f = v;
        }
    }
    __syncthreads();
    for (int pr = threadIdx.x; pr < n_probes; pr += blockDim.x) {
        const PointProbe q = probes[pr];
        const uint64_t c = *q.count;
        if (c < q.capacity) {
            double* o = q.out + 6 * c;
#pragma unroll
            for (int d = 0; d < 6; ++d) o[d] = (double)F[(size_t)d * N + q.idx];
        }
        *q.count = c + 1;
    }
}

constexpr int EN_THREADS = 256;
constexpr int EN_BLOCKS = 1024;

// Field3D::energy (src/arrays/field.rs:219-221): sum of v*v in float64.
__global__ __launch_bounds__(EN_THREADS) void k_energy(const float* __restrict__ F, size_t N,
                                                       double* __restrict__ partials) {
    __shared__ double se[EN_THREADS], sh[EN_THREADS];
    double ae = 0.0, ah = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double e = (double)F[(size_t)d * N + i];
            const double h = (double)F[(size_t)(d + 3) * N + i];
            ae += e * e;
            ah += h * h;
        }
    }
    se[threadIdx.x] = ae;
    sh[threadIdx.x] = ah;
    __syncthreads();
    for (int s = EN_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            se[threadIdx.x] += se[threadIdx.x + s];
            sh[threadIdx.x] += sh[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = se[0];
        partials[2 * blockIdx.x + 1] = sh[0];
    }
}

__global__ __launch_bounds__(EN_THREADS) void k_energy_final(const double* __restrict__ partials, int nb,
                                                             double* __restrict__ out) {
    __shared__ double se[EN_THREADS], sh[EN_THREADS];
    double ae = 0.0, ah = 0.0;
    for (int b = threadIdx.x; b < nb; b += EN_THREADS) {
        ae += partials[2 * b];
        ah += partials[2 * b + 1];
    }
    se[threadIdx.x] = ae;
    sh[threadIdx.x] = ah;
    __syncthreads();
    for (int s = EN_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            se[threadIdx.x] += se[threadIdx.x + s];
            sh[threadIdx.x] += sh[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = se[0];
        out[1] = sh[0];
    }
}

__device__ __forceinline__ size_t mur_cell(const Grid& g, const MurDev& m, int line, int a, int b) {
    int pos[3];
    pos[m.ny] = line;
    pos[m.nyp] = a;
    pos[m.nypp] = b;
    return ((size_t)pos[0] * g.ny + pos[1]) * g.nz + pos[2];
}

// MurAbc::pre_voltage_update (src/extensions/mur_abc.rs:213-254).
__global__ void k_mur_pre(const float* __restrict__ F, Grid g, const MurDev* __restrict__ faces) {
    const MurDev m = faces[blockIdx.y];
    const int cnt = m.s0 * m.s1;
    const float* A = F + (size_t)m.nyp * g.N;
    const float* B = F + (size_t)m.nypp * g.N;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
        const int a = q / m.s1, b = q % m.s1;
        const size_t pb = mur_cell(g, m, m.line, a, b), ps = mur_cell(g, m, m.shift, a, b);
        m.v_nyp[q] = A[ps] - m.coeff * A[pb];
        m.v_nypp[q] = B[ps] - m.coeff * B[pb];
    }
}

// MurAbc::post_voltage_update (src/extensions/mur_abc.rs:257-291).
__global__ void k_mur_post(const float* __restrict__ F, Grid g, const MurDev* __restrict__ faces) {
    const MurDev m = faces[blockIdx.y];
    const int cnt = m.s0 * m.s1;
    const float* A = F + (size_t)m.nyp * g.N;
    const float* B = F + (size_t)m.nypp * g.N;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
        const int a = q / m.s1, b = q % m.s1;
        const size_t ps = mur_cell(g, m, m.shift, a, b);
        m.v_nyp[q] += m.coeff * A[ps];
        m.v_nypp[q] += m.coeff * B[ps];
    }
}

// MurAbc::apply_to_voltages (src/extensions/mur_abc.rs:294-324), one face
// per launch so that faces sharing an edge overwrite in list order.
__global__ void k_mur_apply(float* __restrict__ F, Grid g, const MurDev* __restrict__ faces, int f) {
    const MurDev m = faces[f];
    const int cnt = m.s0 * m.s1;
    float* A = F + (size_t)m.nyp * g.N;
    float* B = F + (size_t)m.nypp * g.N;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
        const int a = q / m.s1, b = q % m.s1;
        const size_t pb = mur_cell(g, m, m.line, a, b);
        A[pb] = m.v_nyp[q];
        B[pb] = m.v_nypp[q];
    }
}

__device__ __forceinline__ void ade_pre_cell(const float* __restrict__ F, size_t N, const AdeDev& a, int c) {
    const size_t p = a.idx[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) a.Eprev[(size_t)d * a.n + c] = F[(size_t)d * N + p];
}

__global__ void k_ade_pre(const float* __restrict__ F, size_t N, AdeDev a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < a.n) ade_pre_cell(F, N, a, c);
}

// Dispersive post_update_e.  Compat: the reference bodies
// (src/extensions/dispersive.rs:171-196, 510-535, 604-630).  Physical: the
// trapezoidal Drude/Debye and second-order Lorentz schemes, with
// K = Cb * delta = dt / eps_cell / (1 + loss) taken from the cell.
__device__ __forceinline__ void ade_post_cell(float* __restrict__ F, const float* __restrict__ Cf, size_t N,
                                              const AdeDev& a, int c) {
    const size_t p = a.idx[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float* fe = F + (size_t)d * N + p;
        const float e = *fe;
        const size_t s = (size_t)d * a.n + c;
        if (a.mode == 0) {
            const float jo = a.J[s];
            if (a.kind == 0) {
                const float jn = a.a * jo + a.c * e;
                a.J[s] = jn;
                *fe = e - a.b * jn;
            } else if (a.kind == 1) {
                const float jn = a.a * jo + a.b * e;
                a.J[s] = jn;
                *fe = e - a.c * jn;
            } else {
                const float pn = a.a * jo + a.b * e;
                a.J[s] = pn;
                const float dpdt = (pn - jo) / a.dt_t;
                *fe = e - a.c * dpdt;
            }
        } else {
            const float K = (float)((double)Cf[(size_t)(3 + d) * N + p] * (double)a.delta[d]);
            const float en = a.Eprev[s];
            const float jn = a.J[s];
            if (a.kind == 1) {
                const float kj = a.pa, bj = a.pb;
                const float enew = (e - (K / 2.f) * ((1.f + kj) * jn + bj * en)) / (1.f + K * bj / 2.f);
                a.J[s] = kj * jn + bj * (enew + en);
                *fe = enew;
            } else if (a.kind == 0) {
                const float pold = a.Jold[s];
                const float pnew = a.pa * jn + a.pb * pold + a.pc * en;
                a.Jold[s] = jn;
                a.J[s] = pnew;
                *fe = e - K * ((pnew - jn) / a.dtf);
            } else {
                const float aa = a.pa, bb = a.pb, dtt = a.dtf;
                const float enew = (e - (K / dtt) * ((aa - 1.f) * jn + bb * en)) / (1.f + K * bb / dtt);
                a.J[s] = aa * jn + bb * (enew + en);
                *fe = enew;
            }
        }
    }
}

__global__ void k_ade_post(float* __restrict__ F, const float* __restrict__ Cf, size_t N, AdeDev a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < a.n) ade_post_cell(F, Cf, N, a, c);
}

// Several dispersive extensions whose cells are disjoint in one launch each
// for the pre and the post step: blockIdx.y selects the extension (their
// order then does not matter).
__global__ void k_ade_pre_multi(const float* __restrict__ F, size_t N, AdeMulti m) {
    const AdeDev& a = m.a[blockIdx.y];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < a.n) ade_pre_cell(F, N, a, c);
}

__global__ void k_ade_post_multi(float* __restrict__ F, const float* __restrict__ Cf, size_t N, AdeMulti m) {
    const AdeDev& a = m.a[blockIdx.y];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < a.n) ade_post_cell(F, Cf, N, a, c);
}

// A work-group owns 256 signals: one thread per signal gathers the float64
// signal value and records it; then the work-group's threads share the
// (frequency, signal) pairs of its signals, so a few signals with many
// frequencies (a port's V and I) update their sums in parallel, and many
// signals (a recording surface) still update each frequency's sums in
// coalesced rows.  Every sum receives the same one update as with a thread
// per signal.  FP32_KAHAN keeps float32 sums with Kahan compensation; the
// phase is evaluated in float64 in both modes.  The work-group tabulates
// exp(-j 2 pi f t) for the E and H sample times of a chunk of frequencies in
// LDS first.
constexpr int DFT_FCHUNK = 128;

// Signal m of a recording box (DftSurface): the node value of E_a1 / E_a2
// (mean of the two edges meeting at the node, the lower one first) or H_a1
// / H_a2 (mean of the four faces around it: the lower normal offset first,
// within it the lower tangential offset), summed in float64 in the order
// oe_ext_nf2ff_create writes the same terms into the tables.  The face is
// picked with selects over the unrolled face list (no dynamic indexing of
// the argument array).
__device__ __forceinline__ double surface_signal(const float* __restrict__ F, size_t N, const DftSurface& s, int m,
                                                 int& h) {
    int ax = 0, q = 0, lo1 = 0, lo2 = 0, n2 = 1, sig0 = 0;
#pragma unroll
    for (int f = 0; f < MAX_SURF_FACES; ++f)
        if (f < s.n && m >= s.f[f].sig0) {
            ax = s.f[f].ax; q = s.f[f].q; lo1 = s.f[f].lo1; lo2 = s.f[f].lo2; n2 = s.f[f].n2; sig0 = s.f[f].sig0;
        }
    const int r = m - sig0, node = r >> 2, c = r & 3;
    const int i1 = lo1 + node / n2, i2 = lo2 + node % n2;
    const long long sx = (long long)s.ny * s.nz, sy = s.nz;
    // strides of ax, a1 = (ax + 1) % 3, a2 = (ax + 2) % 3, and the field
    // components a1, a2
    const long long S0 = ax == 0 ? sx : (ax == 1 ? sy : 1);
    const long long S1 = ax == 0 ? sy : (ax == 1 ? 1 : sx);
    const long long S2 = ax == 0 ? 1 : (ax == 1 ? sx : sy);
    const int a1 = ax == 2 ? 0 : ax + 1, a2 = ax == 0 ? 2 : ax - 1;
    const long long base = q * S0 + i1 * S1 + i2 * S2;
    double v = 0.0;
    if (c < 2) {  // E_a1 / E_a2: the edges at -1 and 0 along that axis
        const long long St = c == 0 ? S1 : S2;
        const float* Fe = F + (size_t)(c == 0 ? a1 : a2) * N;
        v += 0.5 * (double)Fe[base - St];
        v += 0.5 * (double)Fe[base];
        h = 0;
    } else {  // H_a1 (other tangential axis a2) / H_a2 (other a1)
        const long long So = c == 2 ? S2 : S1;
        const float* Fh = F + (size_t)(3 + (c == 2 ? a1 : a2)) * N;
        v += 0.25 * (double)Fh[base - S0 - So];
        v += 0.25 * (double)Fh[base - S0];
        v += 0.25 * (double)Fh[base - So];
        v += 0.25 * (double)Fh[base];
        h = 1;
    }
    return v;
}

// With step_dev the step comes from the device counter (a captured graph):
// sample = sample + st, tE = (t0 + st + 1) dt, tH = (t0 + st + 1/2) dt, the
// same doubles the host forms from the absolute step.  SURF: the signals of
// a recording box from its face description (d.surf) instead of the tables.
template <bool SURF>
__global__ __launch_bounds__(256) void k_dft(const float* __restrict__ F, size_t N, DftDev d, uint64_t sample,
                                            double tE, double tH, const uint64_t* __restrict__ step_dev, double t0,
                                            double dt) {
    __shared__ double ph[2][DFT_FCHUNK][2];  // [E, H][f][cos, -sin]
    __shared__ double sv[256];
    __shared__ uint8_t hv[256];
    if (step_dev) {
        const uint64_t st = *step_dev;
        sample += st;
        const double cur = t0 + (double)st;
        tE = (cur + 1.0) * dt;
        tH = (cur + 0.5) * dt;
    }
    const int m0 = blockIdx.x * 256;
    const int nm = min(256, d.n_sig - m0);
    const int t = threadIdx.x;
    if (t < nm) {
        const int m = m0 + t;
        double s = 0.0;
        int h = 0;
        if constexpr (SURF) {
            s = surface_signal(F, N, d.surf, m, h);
        } else {
            for (uint64_t k = d.off[m]; k < d.off[m + 1]; ++k) s += d.w[k] * (double)F[(size_t)d.comp[k] * N + d.cell[k]];
            h = d.is_h[m] ? 1 : 0;
        }
        if (d.series && sample < d.series_cap) d.series[sample * d.n_sig + m] = s;
        sv[t] = s;
        hv[t] = (uint8_t)h;
    }
    for (int f0 = 0; f0 < d.n_f; f0 += DFT_FCHUNK) {
        const int nf = min(DFT_FCHUNK, d.n_f - f0);
        __syncthreads();  // sv / hv written; the previous chunk's phasors read
        for (int q = t; q < 2 * nf; q += blockDim.x) {
            const int hq = q / nf, f = q - hq * nf;
            double sn, cs;
            sincos(2.0 * M_PI * d.freqs[f0 + f] * (hq ? tH : tE), &sn, &cs);
            ph[hq][f][0] = cs;
            ph[hq][f][1] = -sn;
        }
        __syncthreads();
        for (int q = t; q < nf * nm; q += blockDim.x) {
            const int f = q / nm, g = q - f * nm;
            const double s = sv[g];
            const double cr = ph[hv[g]][f][0], ci = ph[hv[g]][f][1];
            const size_t a = (size_t)(f0 + f) * d.n_sig + m0 + g;
            if (d.prec == 0) {
                d.acc64[2 * a] += s * cr;
                d.acc64[2 * a + 1] += s * ci;
            } else {
                const float svf = (float)s;
                float* acc = d.acc32 + 4 * a;
                const float yr = svf * (float)cr - acc[2];
                const float tr = acc[0] + yr;
                acc[2] = (tr - acc[0]) - yr;
                acc[0] = tr;
                const float yi = svf * (float)ci - acc[3];
                const float ti = acc[1] + yi;
                acc[3] = (ti - acc[1]) - yi;
                acc[1] = ti;
            }
        }
    }
}

// ---------------------------------------------------------------- sweep
// The step index of a sweep lives on the device, so the kernels of one
// step can be captured once in a graph and replayed for every step.

// One thread per structure applies that structure's excitations in list order
// (the order of k_point_ops), with the sample of the current step.
// The sweep's per-step point work, one work-group per structure: its
// excitations (BasicEngine::apply_excitations, src/fdtd/basic_engine.rs:
// 313-395: one thread per source when the structure's source cells are
// distinct, else one thread in list order), then -- after the barrier, so the
// sources are in the fields it samples -- its port DFT (same arithmetic as
// k_dft: signal values to LDS, the phasors of a chunk of frequencies too,
// then the (frequency, signal) accumulators).
__device__ void sweep_dft_body(const SweepDft& x, uint64_t st, double dt) {
    __shared__ double sv[256];
    __shared__ uint8_t hv[256];
    __shared__ double ph[2][DFT_FCHUNK][2];
    const DftDev& d = x.d;
    const uint64_t sample = x.sample0 + st;
    const double cur = x.t0 + (double)st;
    const double tE = (cur + 1.0) * dt, tH = (cur + 0.5) * dt;
    const int m = threadIdx.x;
    if (m < d.n_sig) {
        double s = 0.0;
        for (uint64_t t = d.off[m]; t < d.off[m + 1]; ++t) s += d.w[t] * (double)x.F[(size_t)d.comp[t] * x.N + d.cell[t]];
        if (d.series && sample < d.series_cap) d.series[sample * d.n_sig + m] = s;
        sv[m] = s;
        hv[m] = d.is_h[m] ? 1 : 0;
    }
    for (int f0 = 0; f0 < d.n_f; f0 += DFT_FCHUNK) {
        const int nf = min(DFT_FCHUNK, d.n_f - f0);
        __syncthreads();
        for (int q = threadIdx.x; q < 2 * nf; q += blockDim.x) {
            const int hq = q / nf, f = q - hq * nf;
            double sn, cs;
            sincos(2.0 * M_PI * d.freqs[f0 + f] * (hq ? tH : tE), &sn, &cs);
            ph[hq][f][0] = cs;
            ph[hq][f][1] = -sn;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < nf * d.n_sig; q += blockDim.x) {
            const int f = q / d.n_sig, g = q - f * d.n_sig;
            const double s = sv[g];
            const double cr = ph[hv[g]][f][0], ci = ph[hv[g]][f][1];
            const size_t a = (size_t)(f0 + f) * d.n_sig + g;
            if (d.prec == 0) {
                d.acc64[2 * a] += s * cr;
                d.acc64[2 * a + 1] += s * ci;
            } else {
                const float svf = (float)s;
                float* acc = d.acc32 + 4 * a;
                const float yr = svf * (float)cr - acc[2];
                const float tr = acc[0] + yr;
                acc[2] = (tr - acc[0]) - yr;
                acc[0] = tr;
                const float yi = svf * (float)ci - acc[3];
                const float ti = acc[1] + yi;
                acc[3] = (ti - acc[1]) - yi;
                acc[1] = ti;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_sweep_point(const SweepExc* __restrict__ e, const int* __restrict__ first,
                                                     const int* __restrict__ serial,
                                                     const SweepDft* __restrict__ dfts,
                                                     const int* __restrict__ dft_of, const uint64_t* __restrict__ step,
                                                     double dt) {
    const int s = blockIdx.x;
    const uint64_t st = *step;
    const int q0 = first[s], q1 = first[s + 1];
    if (serial[s]) {
        if (threadIdx.x == 0)
            for (int q = q0; q < q1; ++q) {
                const float v = st < e[q].n_samples ? e[q].samples[st] : 0.f;
                if (e[q].soft) *e[q].target += v;
                else *e[q].target = v;
            }
    } else {
        for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x) {
            const float v = st < e[q].n_samples ? e[q].samples[st] : 0.f;
            if (e[q].soft) *e[q].target += v;
            else *e[q].target = v;
        }
    }
    if (dft_of[s] < 0) return;  // uniform over the work-group
    __syncthreads();
    sweep_dft_body(dfts[dft_of[s]], st, dt);
}

__global__ void k_sweep_step(uint64_t* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// Per-structure energy with the partition of k_energy (EN_BLOCKS blocks per
// structure, the same strides), so every structure's sum equals a
// single-grid engine's.
__global__ __launch_bounds__(EN_THREADS) void k_sweep_energy(const SweepStruct* __restrict__ st,
                                                             const int* __restrict__ ids,
                                                             double* __restrict__ partials) {
    __shared__ double se[EN_THREADS], sh[EN_THREADS];
    const SweepStruct& S = st[ids[blockIdx.y]];
    const size_t N = S.g.N;
    const float* F = S.F;
    double ae = 0.0, ah = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double e = (double)F[(size_t)d * N + i];
            const double h = (double)F[(size_t)(d + 3) * N + i];
            ae += e * e;
            ah += h * h;
        }
    }
    se[threadIdx.x] = ae;
    sh[threadIdx.x] = ah;
    __syncthreads();
    for (int s = EN_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            se[threadIdx.x] += se[threadIdx.x + s];
            sh[threadIdx.x] += sh[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* p = partials + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
        p[0] = se[0];
        p[1] = sh[0];
    }
}

// ----------------------------------------------------------------- UPML
// Upml::pre/post_voltage_update and pre/post_current_update
// (src/extensions/pml.rs:500-657), one thread per region cell, the three
// components in order:
//   pre:  f = c F - cfo flux;  F <- flux;  flux <- f
//   post: f = flux;  flux <- F;  F <- f + cfn F
__device__ __forceinline__ bool upml_cell(const Grid& g, const UpmlDev& u, long q, size_t* p) {
    int r = 0;
#pragma unroll
    for (int t = 1; t < 6; ++t)
        if (t < u.n_reg && q >= u.first[t]) r = t;
    const long l = q - u.first[r];
    const int sz1 = u.size[r][1], sz2 = u.size[r][2];
    const int lk = (int)(l % sz2);
    const long t = l / sz2;
    const int lj = (int)(t % sz1), li = (int)(t / sz1);
    *p = ((size_t)(li + u.start[r][0]) * g.ny + (lj + u.start[r][1])) * g.nz + (lk + u.start[r][2]);
    return true;
}

__global__ void k_upml_pre(float* __restrict__ F, Grid g, UpmlDev u, int e) {
    const long total = u.first[u.n_reg];
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        size_t p;
        upml_cell(g, u, q, &p);
        const int kc = e ? 0 : 3, kfo = e ? 1 : 4, kfl = e ? 6 : 7;
        for (int c = 0; c < 3; ++c) {
            float* f = F + (size_t)(e ? c : 3 + c) * g.N + p;
            float* fl = u.a + ((size_t)kfl * 3 + c) * total + q;
            const float cc = u.a[((size_t)kc * 3 + c) * total + q];
            const float cfo = u.a[((size_t)kfo * 3 + c) * total + q];
            const float old = *f, fo = *fl;
            const float fh = cc * old - cfo * fo;
            *f = fo;
            *fl = fh;
        }
    }
}

__global__ void k_upml_post(float* __restrict__ F, Grid g, UpmlDev u, int e) {
    const long total = u.first[u.n_reg];
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        size_t p;
        upml_cell(g, u, q, &p);
        const int kfn = e ? 2 : 5, kfl = e ? 6 : 7;
        for (int c = 0; c < 3; ++c) {
            float* f = F + (size_t)(e ? c : 3 + c) * g.N + p;
            float* fl = u.a + ((size_t)kfl * 3 + c) * total + q;
            const float cfn = u.a[((size_t)kfn * 3 + c) * total + q];
            const float fh = *fl, fn = *f;
            *fl = fn;
            *f = fh + cfn * fn;
        }
    }
}

int grid_for(size_t n, int threads, int cap) {
    size_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > (size_t)cap) b = cap;
    return (int)b;
}

}  // namespace

hipError_t launch_fill(float* p, size_t n, float v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

int boxes_fill_blocks(const BoxFill& o) {
    const size_t cnt = (size_t)(o.box[1] - o.box[0]) * (o.box[3] - o.box[2]) * (o.box[5] - o.box[4]);
    return (int)((cnt + BF_CELLS - 1) / BF_CELLS);
}

hipError_t launch_boxes_fill(float* Cf, const Grid& g, const BoxFill* ops, const int* wg_first, int n_ops, int n_wg,
                             hipStream_t s) {
    if (n_ops <= 0 || n_wg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_boxes_fill, dim3((unsigned)n_wg), dim3(256), 0, s, Cf, g.N, g.ny, g.nz, ops, wg_first,
                       n_ops);
    return hipGetLastError();
}

hipError_t launch_point_ops(float* F, size_t N, const PointExc* exc, int n_exc, uint64_t step,
                            const PointProbe* probes, int n_probes, hipStream_t s, const uint64_t* step_dev) {
    hipLaunchKernelGGL(k_point_ops, dim3(1), dim3(64), 0, s, F, N, exc, n_exc, step, probes, n_probes, step_dev);
    return hipGetLastError();
}

int energy_blocks() { return EN_BLOCKS; }

hipError_t launch_energy(const float* F, size_t N, double* partials, int blocks, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_energy, dim3(blocks), dim3(EN_THREADS), 0, s, F, N, partials);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(EN_THREADS), 0, s, partials, blocks, out);
    return hipGetLastError();
}

hipError_t launch_mur_pre(float* F, const Grid& g, const MurDev* faces, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mur_pre, dim3(256, n), dim3(256), 0, s, F, g, faces);
    return hipGetLastError();
}

hipError_t launch_mur_post(float* F, const Grid& g, const MurDev* faces, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mur_post, dim3(256, n), dim3(256), 0, s, F, g, faces);
    for (int f = 0; f < n; ++f) hipLaunchKernelGGL(k_mur_apply, dim3(256), dim3(256), 0, s, F, g, faces, f);
    return hipGetLastError();
}

hipError_t launch_sweep_point(const SweepExc* e, const int* first, const int* serial, const SweepDft* dfts,
                              const int* dft_of, int n_struct, const uint64_t* step, double dt, hipStream_t s) {
    if (n_struct <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sweep_point, dim3(n_struct), dim3(256), 0, s, e, first, serial, dfts, dft_of, step, dt);
    return hipGetLastError();
}

hipError_t launch_sweep_step(uint64_t* step, hipStream_t s) {
    hipLaunchKernelGGL(k_sweep_step, dim3(1), dim3(64), 0, s, step);
    return hipGetLastError();
}

hipError_t launch_sweep_energy(const SweepStruct* st, const int* ids, int n, double* partials, double* out,
                               hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sweep_energy, dim3(EN_BLOCKS, n), dim3(EN_THREADS), 0, s, st, ids, partials);
    for (int q = 0; q < n; ++q)
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(EN_THREADS), 0, s, partials + (size_t)q * 2 * EN_BLOCKS,
                           EN_BLOCKS, out + 2 * q);
    return hipGetLastError();
}

hipError_t launch_upml_pre(float* F, const Grid& g, const UpmlDev& u, bool e, hipStream_t s) {
    const long total = u.first[u.n_reg];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_upml_pre, dim3(grid_for((size_t)total, 256, 16384)), dim3(256), 0, s, F, g, u, e ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_upml_post(float* F, const Grid& g, const UpmlDev& u, bool e, hipStream_t s) {
    const long total = u.first[u.n_reg];
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_upml_post, dim3(grid_for((size_t)total, 256, 16384)), dim3(256), 0, s, F, g, u, e ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_dft(const float* F, size_t N, const DftDev& d, uint64_t sample, double tE, double tH,
                      hipStream_t s) {
    if (d.n_sig <= 0) return hipSuccess;
    if (d.surf.n > 0)
        hipLaunchKernelGGL(k_dft<true>, dim3((d.n_sig + 255) / 256), dim3(256), 0, s, F, N, d, sample, tE, tH,
                           (const uint64_t*)nullptr, 0.0, 0.0);
    else
        hipLaunchKernelGGL(k_dft<false>, dim3((d.n_sig + 255) / 256), dim3(256), 0, s, F, N, d, sample, tE, tH,
                           (const uint64_t*)nullptr, 0.0, 0.0);
    return hipGetLastError();
}

hipError_t launch_dft_dev(const float* F, size_t N, const DftDev& d, const uint64_t* step_dev, uint64_t sample0,
                          double t0, double dt, hipStream_t s) {
    if (d.n_sig <= 0) return hipSuccess;
    if (d.surf.n > 0)
        hipLaunchKernelGGL(k_dft<true>, dim3((d.n_sig + 255) / 256), dim3(256), 0, s, F, N, d, sample0, 0.0, 0.0,
                           step_dev, t0, dt);
    else
        hipLaunchKernelGGL(k_dft<false>, dim3((d.n_sig + 255) / 256), dim3(256), 0, s, F, N, d, sample0, 0.0, 0.0,
                           step_dev, t0, dt);
    return hipGetLastError();
}

hipError_t launch_ade_multi(float* F, const float* Cf, size_t N, const AdeMulti& m, bool pre, hipStream_t s) {
    if (m.n <= 0 || m.n > MAX_MERGE_ADE) return hipErrorInvalidValue;
    int cells = 0;
    for (int q = 0; q < m.n; ++q) cells = std::max(cells, m.a[q].n);
    if (cells == 0) return hipSuccess;
    const dim3 grid((cells + 255) / 256, m.n);
    if (pre) hipLaunchKernelGGL(k_ade_pre_multi, grid, dim3(256), 0, s, F, N, m);
    else hipLaunchKernelGGL(k_ade_post_multi, grid, dim3(256), 0, s, F, Cf, N, m);
    return hipGetLastError();
}

hipError_t launch_ade_pre(const float* F, size_t N, const AdeDev& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ade_pre, dim3((a.n + 255) / 256), dim3(256), 0, s, F, N, a);
    return hipGetLastError();
}

hipError_t launch_ade_post(float* F, const float* Cf, size_t N, const AdeDev& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ade_post, dim3((a.n + 255) / 256), dim3(256), 0, s, F, Cf, N, a);
    return hipGetLastError();
}

}  // namespace oe
