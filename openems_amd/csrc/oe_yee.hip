// oe_yee.hip -- FP32 Yee H and E updates for gfx950, with the CPML psi
// recursion and correction fused into the same pass.
//
// Semantics follow the reference engine exactly (src/fdtd/basic_engine.rs):
//   H update :191-254  forward differences, a difference past the last index
//                      is 0; every cell is written.
//   E update :262-306  backward differences, only i, j, k >= 1 are written
//                      (the index-0 planes keep their value).
//   CPML     src/extensions/cpml.rs:443-738 psi recursions over the face
//            slabs; the field correction (physical mode) uses the same
//            differences as the bulk update, so fusing it changes no value.
// Arithmetic is not contracted (a*b + c*d is two roundings and an add, as the
// Rust reference computes it), so results are bit-identical to the oracle.
//
// Work decomposition (CDNA4, wave64): a 256-thread work-group covers bx lanes
// along k (each lane owns V = 4 consecutive k through 16-byte accesses, or
// V = 1) times by rows along j, and marches along i over ci planes.  The
// marched components (Ey, Ez for H; Hy, Hz for E) are carried in registers
// from one plane to the next, so each field element is fetched from HBM once
// per half-step; the j +- 1 and k +- 1 neighbours are re-reads that hit L1/L2.
// Consecutive tiles are remapped onto one XCD so that the j-halo rows shared
// by neighbouring work-groups stay in that XCD's L2.

#include "oe_internal.h"

#pragma clang fp contract(off)

namespace oe {
namespace {

template <int V>
struct Vec {
    float v[V];
};

template <int V>
__device__ __forceinline__ Vec<V> ldv(const float* __restrict__ p) {
    Vec<V> r;
    if constexpr (V == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p);
        r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
    } else {
        r.v[0] = *p;
    }
    return r;
}

template <int V>
__device__ __forceinline__ void stv(float* __restrict__ p, const Vec<V>& a) {
    if constexpr (V == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]);
    } else {
        *p = a.v[0];
    }
}

template <int V>
__device__ __forceinline__ Vec<V> zerov() {
    Vec<V> r;
#pragma unroll
    for (int e = 0; e < V; ++e) r.v[e] = 0.f;
    return r;
}

struct Tile {
    int k0, j, i0, i1;
};

// Block id -> (k-block, j-block, i-chunk).  Hardware hands consecutive
// work-group ids to the 8 XCDs round-robin; the remap gives XCD x the
// contiguous logical range [x*per, (x+1)*per) so that j-neighbouring tiles
// run on the same XCD at the same time.
template <int V>
__device__ __forceinline__ Tile decode(const YeeArgs& a) {
    int bid = blockIdx.x;
    if (a.xcd_remap) {
        const int per = gridDim.x >> 3;
        if (bid < (per << 3)) bid = (bid & 7) * per + (bid >> 3);
    }
    const int kb = bid % a.n_kb;
    const int t = bid / a.n_kb;
    const int jb = t % a.n_jb;
    const int ib = t / a.n_jb;
    const int lane = threadIdx.x;
    const int kx = lane % a.t.bx;
    const int jy = lane / a.t.bx;
    Tile r;
    r.k0 = (kb * a.t.bx + kx) * V;
    r.j = jb * a.t.by + jy;
    r.i0 = ib * a.t.ci;
    r.i1 = min(r.i0 + a.t.ci, a.g.nx);
    return r;
}

__device__ __forceinline__ float tab(const CpmlArgs& c, int f, int t, int l) {
    return c.tab[(f * CPML_TABLES + t) * c.Lmax + l];
}

template <int V, bool CP>
__global__ __launch_bounds__(256) void k_update_h(YeeArgs a, CpmlArgs c) {
    const Tile t = decode<V>(a);
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    if (t.k0 >= nz || t.j >= ny) return;
    const size_t N = a.g.N;
    const size_t plane = (size_t)ny * nz;
    const float* __restrict__ Ex = a.F;
    const float* __restrict__ Ey = a.F + N;
    const float* __restrict__ Ez = a.F + 2 * N;
    float* __restrict__ Hx = a.F + 3 * N;
    float* __restrict__ Hy = a.F + 4 * N;
    float* __restrict__ Hz = a.F + 5 * N;
    const float* __restrict__ da0 = a.Cf + 6 * N;
    const float* __restrict__ da1 = a.Cf + 7 * N;
    const float* __restrict__ da2 = a.Cf + 8 * N;
    const float* __restrict__ db0 = a.Cf + 9 * N;
    const float* __restrict__ db1 = a.Cf + 10 * N;
    const float* __restrict__ db2 = a.Cf + 11 * N;

    const int j = t.j, k0 = t.k0;
    const bool hj = j + 1 < ny;
    const bool hk_last = k0 + V < nz;
    size_t p = (size_t)t.i0 * plane + (size_t)j * nz + k0;
    Vec<V> ey_c = ldv<V>(Ey + p);
    Vec<V> ez_c = ldv<V>(Ez + p);

    for (int i = t.i0; i < t.i1; ++i, p += plane) {
        const bool hi = i + 1 < nx;
        Vec<V> ey_n = ey_c, ez_n = ez_c;
        if (hi) {
            ey_n = ldv<V>(Ey + p + plane);
            ez_n = ldv<V>(Ez + p + plane);
        }
        const Vec<V> ex_c = ldv<V>(Ex + p);
        Vec<V> ez_j = ez_c, ex_j = ex_c;
        if (hj) {
            ez_j = ldv<V>(Ez + p + nz);
            ex_j = ldv<V>(Ex + p + nz);
        }
        float ey_kl = 0.f, ex_kl = 0.f;
        if (hk_last) {
            ey_kl = Ey[p + V];
            ex_kl = Ex[p + V];
        }
        Vec<V> hx = ldv<V>(Hx + p), hy = ldv<V>(Hy + p), hz = ldv<V>(Hz + p);
        const Vec<V> a0 = ldv<V>(da0 + p), a1 = ldv<V>(da1 + p), a2 = ldv<V>(da2 + p);
        const Vec<V> b0 = ldv<V>(db0 + p), b1 = ldv<V>(db1 + p), b2 = ldv<V>(db2 + p);

        // Differences kept for the CPML pass.
        float dEz_x[V], dEy_x[V], dEx_y[V], dEz_y[V], dEy_z[V], dEx_z[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const bool hk = (e < V - 1) || hk_last;
            const float eyk = (e < V - 1) ? ey_c.v[(e < V - 1) ? e + 1 : e] : ey_kl;
            const float exk = (e < V - 1) ? ex_c.v[(e < V - 1) ? e + 1 : e] : ex_kl;
            const float dez_dy = hj ? ez_j.v[e] - ez_c.v[e] : 0.f;
            const float dey_dz = hk ? eyk - ey_c.v[e] : 0.f;
            const float cx = dez_dy - dey_dz;
            hx.v[e] = a0.v[e] * hx.v[e] + b0.v[e] * cx;
            const float dex_dz = hk ? exk - ex_c.v[e] : 0.f;
            const float dez_dx = hi ? ez_n.v[e] - ez_c.v[e] : 0.f;
            const float cy = dex_dz - dez_dx;
            hy.v[e] = a1.v[e] * hy.v[e] + b1.v[e] * cy;
            const float dey_dx = hi ? ey_n.v[e] - ey_c.v[e] : 0.f;
            const float dex_dy = hj ? ex_j.v[e] - ex_c.v[e] : 0.f;
            const float cz = dey_dx - dex_dy;
            hz.v[e] = a2.v[e] * hz.v[e] + b2.v[e] * cz;
            dEz_x[e] = dez_dx; dEy_x[e] = dey_dx;
            dEx_y[e] = dex_dy; dEz_y[e] = dez_dy;
            dEy_z[e] = dey_dz; dEx_z[e] = dex_dz;
        }

        if constexpr (CP) {
            const bool phys = c.physical != 0;
            // x faces (psi_pos <- dEz/dx for Hy, psi_neg <- dEy/dx for Hz)
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                if (!(c.on_mask & (1 << f))) continue;
                const int l = i - c.g0[f];
                if (l < 0 || l >= c.L[f]) continue;
                const size_t q = ((size_t)l * ny + j) * nz + k0;
                const float bb = tab(c, f, TAB_BH, l), aa = tab(c, f, TAB_AH, l), km = tab(c, f, TAB_KH, l);
                Vec<V> pp = ldv<V>(c.ph_pos[f] + q), pn = ldv<V>(c.ph_neg[f] + q);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    pp.v[e] = bb * pp.v[e] + aa * dEz_x[e];
                    pn.v[e] = bb * pn.v[e] + aa * dEy_x[e];
                    if (phys) {
                        hy.v[e] = hy.v[e] + b1.v[e] * (pp.v[e] - km * dEz_x[e]);
                        hz.v[e] = hz.v[e] + b2.v[e] * (km * dEy_x[e] - pn.v[e]);
                    }
                }
                stv<V>(c.ph_pos[f] + q, pp);
                stv<V>(c.ph_neg[f] + q, pn);
            }
            // y faces (psi_pos <- dEx/dy for Hz, psi_neg <- dEz/dy for Hx)
#pragma unroll
            for (int f = 2; f < 4; ++f) {
                if (!(c.on_mask & (1 << f))) continue;
                const int l = j - c.g0[f];
                if (l < 0 || l >= c.L[f]) continue;
                const size_t q = ((size_t)i * c.L[f] + l) * nz + k0;
                const float bb = tab(c, f, TAB_BH, l), aa = tab(c, f, TAB_AH, l), km = tab(c, f, TAB_KH, l);
                Vec<V> pp = ldv<V>(c.ph_pos[f] + q), pn = ldv<V>(c.ph_neg[f] + q);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    pp.v[e] = bb * pp.v[e] + aa * dEx_y[e];
                    pn.v[e] = bb * pn.v[e] + aa * dEz_y[e];
                    if (phys) {
                        hz.v[e] = hz.v[e] + b2.v[e] * (pp.v[e] - km * dEx_y[e]);
                        hx.v[e] = hx.v[e] + b0.v[e] * (km * dEz_y[e] - pn.v[e]);
                    }
                }
                stv<V>(c.ph_pos[f] + q, pp);
                stv<V>(c.ph_neg[f] + q, pn);
            }
            // z faces (psi_pos <- dEy/dz for Hx, psi_neg <- dEx/dz for Hy)
#pragma unroll
            for (int f = 4; f < 6; ++f) {
                if (!(c.on_mask & (1 << f))) continue;
                const int L = c.L[f];
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const int l = k0 + e - c.g0[f];
                    if (l < 0 || l >= L) continue;
                    const size_t q = ((size_t)i * ny + j) * L + l;
                    const float bb = tab(c, f, TAB_BH, l), aa = tab(c, f, TAB_AH, l), km = tab(c, f, TAB_KH, l);
                    const float pp = bb * c.ph_pos[f][q] + aa * dEy_z[e];
                    const float pn = bb * c.ph_neg[f][q] + aa * dEx_z[e];
                    c.ph_pos[f][q] = pp;
                    c.ph_neg[f][q] = pn;
                    if (phys) {
                        hx.v[e] = hx.v[e] + b0.v[e] * (pp - km * dEy_z[e]);
                        hy.v[e] = hy.v[e] + b1.v[e] * (km * dEx_z[e] - pn);
                    }
                }
            }
        }

        stv<V>(Hx + p, hx);
        stv<V>(Hy + p, hy);
        stv<V>(Hz + p, hz);
        ey_c = ey_n;
        ez_c = ez_n;
    }
}

template <int V, bool CP>
__global__ __launch_bounds__(256) void k_update_e(YeeArgs a, CpmlArgs c) {
    const Tile t = decode<V>(a);
    const int ny = a.g.ny, nz = a.g.nz;
    if (t.k0 >= nz || t.j >= ny) return;
    const size_t N = a.g.N;
    const size_t plane = (size_t)ny * nz;
    float* __restrict__ Ex = a.F;
    float* __restrict__ Ey = a.F + N;
    float* __restrict__ Ez = a.F + 2 * N;
    const float* __restrict__ Hx = a.F + 3 * N;
    const float* __restrict__ Hy = a.F + 4 * N;
    const float* __restrict__ Hz = a.F + 5 * N;
    const float* __restrict__ ca0 = a.Cf;
    const float* __restrict__ ca1 = a.Cf + N;
    const float* __restrict__ ca2 = a.Cf + 2 * N;
    const float* __restrict__ cb0 = a.Cf + 3 * N;
    const float* __restrict__ cb1 = a.Cf + 4 * N;
    const float* __restrict__ cb2 = a.Cf + 5 * N;

    const int j = t.j, k0 = t.k0;
    const bool gj = j > 0;
    const bool gk_first = k0 > 0;
    size_t p = (size_t)t.i0 * plane + (size_t)j * nz + k0;
    Vec<V> hy_p = zerov<V>(), hz_p = zerov<V>();
    if (t.i0 > 0) {
        hy_p = ldv<V>(Hy + p - plane);
        hz_p = ldv<V>(Hz + p - plane);
    }

    for (int i = t.i0; i < t.i1; ++i, p += plane) {
        const bool gi = i > 0;
        const Vec<V> hx_c = ldv<V>(Hx + p), hy_c = ldv<V>(Hy + p), hz_c = ldv<V>(Hz + p);
        Vec<V> hz_j = hz_c, hx_j = hx_c;
        if (gj) {
            hz_j = ldv<V>(Hz + p - nz);
            hx_j = ldv<V>(Hx + p - nz);
        }
        float hy_kf = 0.f, hx_kf = 0.f;
        if (gk_first) {
            hy_kf = Hy[p - 1];
            hx_kf = Hx[p - 1];
        }
        float dHz_y[V], dHy_z[V], dHx_z[V], dHz_x[V], dHy_x[V], dHx_y[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const bool gk = (e > 0) || gk_first;
            const float hyk = (e > 0) ? hy_c.v[(e > 0) ? e - 1 : e] : hy_kf;
            const float hxk = (e > 0) ? hx_c.v[(e > 0) ? e - 1 : e] : hx_kf;
            dHz_y[e] = gj ? hz_c.v[e] - hz_j.v[e] : 0.f;
            dHy_z[e] = gk ? hy_c.v[e] - hyk : 0.f;
            dHx_z[e] = gk ? hx_c.v[e] - hxk : 0.f;
            dHz_x[e] = gi ? hz_c.v[e] - hz_p.v[e] : 0.f;
            dHy_x[e] = gi ? hy_c.v[e] - hy_p.v[e] : 0.f;
            dHx_y[e] = gj ? hx_c.v[e] - hx_j.v[e] : 0.f;
        }

        const bool row = gi && gj;
        Vec<V> ex, ey, ez, c0, c1, c2;
        if (row) {
            ex = ldv<V>(Ex + p); ey = ldv<V>(Ey + p); ez = ldv<V>(Ez + p);
            const Vec<V> s0 = ldv<V>(ca0 + p), s1 = ldv<V>(ca1 + p), s2 = ldv<V>(ca2 + p);
            c0 = ldv<V>(cb0 + p); c1 = ldv<V>(cb1 + p); c2 = ldv<V>(cb2 + p);
#pragma unroll
            for (int e = 0; e < V; ++e) {
                if (!((e > 0) || gk_first)) continue;  // k == 0: not written
                const float cx = dHz_y[e] - dHy_z[e];
                ex.v[e] = s0.v[e] * ex.v[e] + c0.v[e] * cx;
                const float cy = dHx_z[e] - dHz_x[e];
                ey.v[e] = s1.v[e] * ey.v[e] + c1.v[e] * cy;
                const float cz = dHy_x[e] - dHx_y[e];
                ez.v[e] = s2.v[e] * ez.v[e] + c2.v[e] * cz;
            }
        }

        if constexpr (CP) {
            const bool phys = (c.physical != 0) && row;
            // x faces (psi_pos <- dHz/dx for Ey, psi_neg <- dHy/dx for Ez)
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                if (!(c.on_mask & (1 << f))) continue;
                const int l = i - c.g0[f];
                if (l < 0 || l >= c.L[f]) continue;
                const size_t q = ((size_t)l * ny + j) * nz + k0;
                const float bb = tab(c, f, TAB_BE, l), aa = tab(c, f, TAB_AE, l), km = tab(c, f, TAB_KE, l);
                Vec<V> pp = ldv<V>(c.pe_pos[f] + q), pn = ldv<V>(c.pe_neg[f] + q);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    pp.v[e] = bb * pp.v[e] + aa * dHz_x[e];
                    pn.v[e] = bb * pn.v[e] + aa * dHy_x[e];
                    if (phys && ((e > 0) || gk_first)) {
                        ey.v[e] = ey.v[e] + c1.v[e] * (pp.v[e] - km * dHz_x[e]);
                        ez.v[e] = ez.v[e] + c2.v[e] * (km * dHy_x[e] - pn.v[e]);
                    }
                }
                stv<V>(c.pe_pos[f] + q, pp);
                stv<V>(c.pe_neg[f] + q, pn);
            }
            // y faces (psi_pos <- dHx/dy for Ez, psi_neg <- dHz/dy for Ex)
#pragma unroll
            for (int f = 2; f < 4; ++f) {
                if (!(c.on_mask & (1 << f))) continue;
                const int l = j - c.g0[f];
                if (l < 0 || l >= c.L[f]) continue;
                const size_t q = ((size_t)i * c.L[f] + l) * nz + k0;
                const float bb = tab(c, f, TAB_BE, l), aa = tab(c, f, TAB_AE, l), km = tab(c, f, TAB_KE, l);
                Vec<V> pp = ldv<V>(c.pe_pos[f] + q), pn = ldv<V>(c.pe_neg[f] + q);
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    pp.v[e] = bb * pp.v[e] + aa * dHx_y[e];
                    pn.v[e] = bb * pn.v[e] + aa * dHz_y[e];
                    if (phys && ((e > 0) || gk_first)) {
                        ez.v[e] = ez.v[e] + c2.v[e] * (pp.v[e] - km * dHx_y[e]);
                        ex.v[e] = ex.v[e] + c0.v[e] * (km * dHz_y[e] - pn.v[e]);
                    }
                }
                stv<V>(c.pe_pos[f] + q, pp);
                stv<V>(c.pe_neg[f] + q, pn);
            }
            // z faces (psi_pos <- dHy/dz for Ex, psi_neg <- dHx/dz for Ey)
#pragma unroll
            for (int f = 4; f < 6; ++f) {
                if (!(c.on_mask & (1 << f))) continue;
                const int L = c.L[f];
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const int l = k0 + e - c.g0[f];
                    if (l < 0 || l >= L) continue;
                    const size_t q = ((size_t)i * ny + j) * L + l;
                    const float bb = tab(c, f, TAB_BE, l), aa = tab(c, f, TAB_AE, l), km = tab(c, f, TAB_KE, l);
                    const float pp = bb * c.pe_pos[f][q] + aa * dHy_z[e];
                    const float pn = bb * c.pe_neg[f][q] + aa * dHx_z[e];
                    c.pe_pos[f][q] = pp;
                    c.pe_neg[f][q] = pn;
                    if (phys && ((e > 0) || gk_first)) {
                        ex.v[e] = ex.v[e] + c0.v[e] * (pp - km * dHy_z[e]);
                        ey.v[e] = ey.v[e] + c1.v[e] * (km * dHx_z[e] - pn);
                    }
                }
            }
        }

        if (row) {
            stv<V>(Ex + p, ex);
            stv<V>(Ey + p, ey);
            stv<V>(Ez + p, ez);
        }
        hy_p = hy_c;
        hz_p = hz_c;
    }
}

template <int V, bool CP>
hipError_t launch_h_t(const YeeArgs& a, const CpmlArgs& c, hipStream_t s) {
    const unsigned nb = (unsigned)a.n_kb * a.n_jb * a.n_ib;
    hipLaunchKernelGGL((k_update_h<V, CP>), dim3(nb), dim3(256), 0, s, a, c);
    return hipGetLastError();
}

template <int V, bool CP>
hipError_t launch_e_t(const YeeArgs& a, const CpmlArgs& c, hipStream_t s) {
    const unsigned nb = (unsigned)a.n_kb * a.n_jb * a.n_ib;
    hipLaunchKernelGGL((k_update_e<V, CP>), dim3(nb), dim3(256), 0, s, a, c);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_update_h(const YeeArgs& a, const CpmlArgs* c, hipStream_t s) {
    CpmlArgs none{};
    if (a.t.vec == 4) return c ? launch_h_t<4, true>(a, *c, s) : launch_h_t<4, false>(a, none, s);
    return c ? launch_h_t<1, true>(a, *c, s) : launch_h_t<1, false>(a, none, s);
}

hipError_t launch_update_e(const YeeArgs& a, const CpmlArgs* c, hipStream_t s) {
    CpmlArgs none{};
    if (a.t.vec == 4) return c ? launch_e_t<4, true>(a, *c, s) : launch_e_t<4, false>(a, none, s);
    return c ? launch_e_t<1, true>(a, *c, s) : launch_e_t<1, false>(a, none, s);
}

}  // namespace oe
