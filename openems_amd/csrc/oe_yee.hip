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
//
// CPML: each cell is in at most one face slab per axis (faces of one axis do
// not overlap), so the face is selected arithmetically and the psi loads are
// issued unconditionally together with the bulk loads -- cells outside every
// slab read a zeroed dummy line -- which keeps the psi latency hidden behind
// the bulk loads instead of serialising it after them.

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
__device__ __forceinline__ void stv(float* __restrict__ p, const Vec<V>& a);

// Streaming (nontemporal) accesses for data without reuse inside a launch:
// the coefficient arrays and the component being updated (load and store).
// On gfx950 the nt hint keeps these streams from displacing the neighbour
// lines (the other field's j/k halos) in L2/MALL.  Measured on the 256^3
// CPML workload: 4.7 -> 5.7 TB/s for the H update, 4.8 -> 5.8 TB/s for E.
#ifndef OE_NT_COEF
#define OE_NT_COEF 1
#endif
#ifndef OE_NT_OWN
#define OE_NT_OWN 1
#endif
#ifndef OE_NT_STORE
#define OE_NT_STORE 1
#endif
typedef float f4v __attribute__((ext_vector_type(4)));

template <int V, bool NT>
__device__ __forceinline__ Vec<V> ldn(const float* __restrict__ p) {
    if constexpr (!NT) {
        return ldv<V>(p);
    } else {
        Vec<V> r;
        if constexpr (V == 4) {
            const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
            r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
        } else {
            r.v[0] = __builtin_nontemporal_load(p);
        }
        return r;
    }
}

template <int V, bool NT>
__device__ __forceinline__ void stn(float* __restrict__ p, const Vec<V>& a) {
    if constexpr (!NT) {
        stv<V>(p, a);
    } else if constexpr (V == 4) {
        f4v t = {a.v[0], a.v[1], a.v[2], a.v[3]};
        __builtin_nontemporal_store(t, reinterpret_cast<f4v*>(p));
    } else {
        __builtin_nontemporal_store(a.v[0], p);
    }
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

// Material table (oe_coef.hip): V row indices of a lane at cell p and the
// rows' (a, b) pairs of x, y, z from LDS.  The index is read from m8 (one
// byte per cell, tables of at most 256 rows) when present, else from m16
// (nontemporal, one 8-byte load for V = 4); neither stands for a table of
// one row, whose indices are all 0 and are not read.
typedef unsigned int u2v __attribute__((ext_vector_type(2)));

struct MatIndex {
    const uint16_t* m16;
    const uint8_t* m8;
};

// A lane's V row indices at cell p, as loaded (unpacked by coef_from): the E
// update loads them one plane ahead.
struct MatIdx {
    unsigned a = 0, b = 0;
};

template <int V>
__device__ __forceinline__ MatIdx coef_index(const MatIndex& mi, size_t p) {
    MatIdx r;
    if (mi.m8 != nullptr) {
        if constexpr (V == 4) r.a = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(mi.m8 + p));
        else r.a = mi.m8[p];
    } else if (mi.m16 != nullptr) {
        if constexpr (V == 4) {
            const u2v t = __builtin_nontemporal_load(reinterpret_cast<const u2v*>(mi.m16 + p));
            r.a = t.x;
            r.b = t.y;
        } else {
            r.a = mi.m16[p];
        }
    }
    return r;
}

template <int V>
__device__ __forceinline__ void coef_from(const MatIndex& mi, const MatIdx& x, const float2* tab, Vec<V>& a0,
                                          Vec<V>& a1, Vec<V>& a2, Vec<V>& b0, Vec<V>& b1, Vec<V>& b2) {
    unsigned m[V];
    if constexpr (V == 4) {
        if (mi.m8 != nullptr) {
            m[0] = x.a & 0xffu; m[1] = (x.a >> 8) & 0xffu; m[2] = (x.a >> 16) & 0xffu; m[3] = x.a >> 24;
        } else {
            m[0] = x.a & 0xffffu; m[1] = x.a >> 16; m[2] = x.b & 0xffffu; m[3] = x.b >> 16;
        }
    } else {
        m[0] = x.a;
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const float2 cx = tab[3 * m[e]], cy = tab[3 * m[e] + 1], cz = tab[3 * m[e] + 2];
        a0.v[e] = cx.x; b0.v[e] = cx.y;
        a1.v[e] = cy.x; b1.v[e] = cy.y;
        a2.v[e] = cz.x; b2.v[e] = cz.y;
    }
}

// The index load and the row reads at once (the H update's form).
template <int V>
__device__ __forceinline__ void coef_rows(const MatIndex& mi, size_t p, const float2* tab, Vec<V>& a0,
                                          Vec<V>& a1, Vec<V>& a2, Vec<V>& b0, Vec<V>& b1, Vec<V>& b2) {
    unsigned m[V];
    if (mi.m8 != nullptr) {
        if constexpr (V == 4) {
            const unsigned t = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(mi.m8 + p));
            m[0] = t & 0xffu; m[1] = (t >> 8) & 0xffu; m[2] = (t >> 16) & 0xffu; m[3] = t >> 24;
        } else {
            m[0] = mi.m8[p];
        }
    } else if (mi.m16 == nullptr) {
#pragma unroll
        for (int e = 0; e < V; ++e) m[e] = 0;
    } else if constexpr (V == 4) {
        const u2v t = __builtin_nontemporal_load(reinterpret_cast<const u2v*>(mi.m16 + p));
        m[0] = t.x & 0xffffu; m[1] = t.x >> 16; m[2] = t.y & 0xffffu; m[3] = t.y >> 16;
    } else {
        m[0] = mi.m16[p];
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const float2 x = tab[3 * m[e]], y = tab[3 * m[e] + 1], z = tab[3 * m[e] + 2];
        a0.v[e] = x.x; b0.v[e] = x.y;
        a1.v[e] = y.x; b1.v[e] = y.y;
        a2.v[e] = z.x; b2.v[e] = z.y;
    }
}

// kfirst / klast: the lane's left / right k neighbour along the row is not
// the adjacent lane of its wave (row ends, wave ends), so the neighbour value
// is loaded instead of taken from that lane.
struct Tile {
    int k0, j, i0, i1, jend;
    bool kfirst, klast;
};

// Block id -> (box, k-block, j-block, i-chunk).  Hardware hands consecutive
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
    int b = 0;
#pragma unroll
    for (int q = 1; q < MAX_BOXES; ++q)
        if (q < a.nbox && bid >= a.first[q]) b = q;
    const Box& B = a.box[b];
    bid -= a.first[b];
    const int kb = bid % a.n_kb;
    const int t = bid / a.n_kb;
    const int jb = t % B.n_jb;
    const int ib = t / B.n_jb;
    const int lane = threadIdx.x;
    const int kx = lane % a.t.bx;
    const int jy = lane / a.t.bx;
    Tile r;
    r.k0 = (kb * a.t.bx + kx) * V;
    r.j = B.j0 + jb * a.t.by + jy;
    r.i0 = B.i0 + ib * a.t.ci;
    r.i1 = min(r.i0 + a.t.ci, B.i1);
    r.jend = B.j1;
    r.kfirst = kx == 0 || (lane & 63) == 0;
    r.klast = kx == a.t.bx - 1 || (lane & 63) == 63;
    return r;
}

// Face of one axis holding global index g, or -1 (faces 2d, 2d+1).
__device__ __forceinline__ int face_of(const CpmlArgs& c, int d, int g, int& l) {
    const int f0 = 2 * d, f1 = 2 * d + 1;
    if ((c.on_mask >> f0 & 1) && g < c.L[f0]) { l = g; return f0; }
    if ((c.on_mask >> f1 & 1) && g >= c.g0[f1]) { l = g - c.g0[f1]; return f1; }
    l = 0;
    return -1;
}

// psi and table operands of one axis for the V cells of a lane.
template <int V>
struct Psi {
    int f;
    const float* p;  // psi_pos address (or dummy)
    const float* n;  // psi_neg address (or dummy)
    float* wp;       // where the new psi_pos / psi_neg go (same offsets)
    float* wn;
    Vec<V> pp, pn;   // loaded values
    Vec<V> b, a, km; // per-cell table values
};

// Loads psi of face f (2d or 2d+1, or -1 for the dummy line) and the three
// tables; tables of x and y faces are uniform over the V cells, those of z
// faces step with k.  Faces are chosen with selects (no dynamic indexing of
// the argument arrays, which would spill them to scratch).
template <int V, bool ZAXIS>
__device__ __forceinline__ void psi_load(Psi<V>& s, const CpmlArgs& c, int d, int f, int l, size_t q,
                                         float* const* pos, float* const* neg, float* const* wpos,
                                         float* const* wneg, int tb, int ta, int tk) {
    s.f = f;
    const bool on = f >= 0;
    const bool lo = f == 2 * d;
    const int fs = on ? f : 2 * d;
    float* P = lo ? pos[2 * d] : pos[2 * d + 1];
    float* Q = lo ? neg[2 * d] : neg[2 * d + 1];
    float* WP = lo ? wpos[2 * d] : wpos[2 * d + 1];
    float* WQ = lo ? wneg[2 * d] : wneg[2 * d + 1];
    s.p = on ? P + q : c.dummy;
    s.n = on ? Q + q : c.dummy;
    s.wp = on ? WP + q : nullptr;
    s.wn = on ? WQ + q : nullptr;
    s.pp = ldv<V>(s.p);
    s.pn = ldv<V>(s.n);
    const float* T = c.tab + (size_t)fs * CPML_TABLES * c.Lmax;
    const int lt = on ? l : 0;
    if constexpr (ZAXIS) {
        s.b = ldv<V>(T + tb * c.Lmax + lt);
        s.a = ldv<V>(T + ta * c.Lmax + lt);
        s.km = ldv<V>(T + tk * c.Lmax + lt);
    } else {
        const float b = T[tb * c.Lmax + lt], a = T[ta * c.Lmax + lt], km = T[tk * c.Lmax + lt];
#pragma unroll
        for (int e = 0; e < V; ++e) { s.b.v[e] = b; s.a.v[e] = a; s.km.v[e] = km; }
    }
}

// psi <- b psi + a D for the pos and neg recursions, then (physical) the
// correction of the Q-term component c1 (+psi) and P-term component c2 (-psi):
//   c1 += C1 (psi_pos - km D_pos),  c2 += C2 (km D_neg - psi_neg)
template <int V, bool PHYS>
__device__ __forceinline__ void psi_apply(Psi<V>& s, const float* dpos, const float* dneg, Vec<V>& c1,
                                          const Vec<V>& C1, Vec<V>& c2, const Vec<V>& C2, const bool* upd) {
#pragma unroll
    for (int e = 0; e < V; ++e) {
        s.pp.v[e] = s.b.v[e] * s.pp.v[e] + s.a.v[e] * dpos[e];
        s.pn.v[e] = s.b.v[e] * s.pn.v[e] + s.a.v[e] * dneg[e];
        if constexpr (PHYS) {
            if (s.f >= 0 && upd[e]) {  // off-slab cells computed psi from the dummy line: no correction
                c1.v[e] = c1.v[e] + C1.v[e] * (s.pp.v[e] - s.km.v[e] * dpos[e]);
                c2.v[e] = c2.v[e] + C2.v[e] * (s.km.v[e] * dneg[e] - s.pn.v[e]);
            }
        }
    }
}

template <int V>
__device__ __forceinline__ void psi_store(const Psi<V>& s) {
    if (s.f >= 0) {
        stv<V>(s.wp, s.pp);
        stv<V>(s.wn, s.pn);
    }
}

// AX: CPML axes compiled in (AX_X | AX_Y | AX_Z); PHYS: apply the field
// correction (else psi only, as the reference); TAB: coefficients from the
// material table (mat, tab in LDS) instead of the arrays in Cf.
template <int V, int AX, bool PHYS, bool TAB>
__device__ __forceinline__ void h_body(const Grid& g, const float* Fin, float* Fout, const float* __restrict__ Cf,
                                       const MatIndex mat, const float2* tab,
                                       const CpmlArgs& c, const Tile& t) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (t.k0 >= nz || t.j >= t.jend) return;
    const size_t N = g.N;
    const size_t plane = (size_t)ny * nz;
    const float* __restrict__ Ex = Fin;
    const float* __restrict__ Ey = Fin + N;
    const float* __restrict__ Ez = Fin + 2 * N;
    // the old H is read from Fin and the new H stored to Fout (the same
    // arrays for an in-place update: no __restrict__ across the pair)
    const float* Hx = Fin + 3 * N;
    const float* Hy = Fin + 4 * N;
    const float* Hz = Fin + 5 * N;
    float* Hxw = Fout + 3 * N;
    float* Hyw = Fout + 4 * N;
    float* Hzw = Fout + 5 * N;
    const float* __restrict__ da0 = Cf + 6 * N;
    const float* __restrict__ da1 = Cf + 7 * N;
    const float* __restrict__ da2 = Cf + 8 * N;
    const float* __restrict__ db0 = Cf + 9 * N;
    const float* __restrict__ db1 = Cf + 10 * N;
    const float* __restrict__ db2 = Cf + 11 * N;

    const int j = t.j, k0 = t.k0;
    const bool hj = j + 1 < ny;
    const bool hk_last = k0 + V < nz;
    int ly = 0, lz = 0;
    int fy = -1, fz = -1;
    if constexpr ((AX & AX_Y) != 0) fy = face_of(c, 1, j, ly);
    if constexpr ((AX & AX_Z) != 0) fz = face_of(c, 2, k0, lz);
    size_t p = (size_t)t.i0 * plane + (size_t)j * nz + k0;
    Vec<V> ey_c = ldv<V>(Ey + p);
    Vec<V> ez_c = ldv<V>(Ez + p);
    bool upd[V];
#pragma unroll
    for (int e = 0; e < V; ++e) upd[e] = true;

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
        // the k + V neighbour: the next lane's first cell, or a load where
        // that lane holds another row (or sits in another wave)
        const float ey_sh = __shfl_down(ey_c.v[0], 1, 64);
        const float ex_sh = __shfl_down(ex_c.v[0], 1, 64);
        float ey_kl = 0.f, ex_kl = 0.f;
        if (hk_last) {
            ey_kl = t.klast ? Ey[p + V] : ey_sh;
            ex_kl = t.klast ? Ex[p + V] : ex_sh;
        }
        Vec<V> hx = ldn<V, OE_NT_OWN>(Hx + p), hy = ldn<V, OE_NT_OWN>(Hy + p), hz = ldn<V, OE_NT_OWN>(Hz + p);
        Vec<V> a0, a1, a2, b0, b1, b2;
        if constexpr (TAB) {
            coef_rows<V>(mat, p, tab, a0, a1, a2, b0, b1, b2);
        } else {
            a0 = ldn<V, OE_NT_COEF>(da0 + p); a1 = ldn<V, OE_NT_COEF>(da1 + p); a2 = ldn<V, OE_NT_COEF>(da2 + p);
            b0 = ldn<V, OE_NT_COEF>(db0 + p); b1 = ldn<V, OE_NT_COEF>(db1 + p); b2 = ldn<V, OE_NT_COEF>(db2 + p);
        }

        Psi<V> px, py, pz;
        if constexpr ((AX & AX_X) != 0) {
            int lx;
            const int fx = face_of(c, 0, i, lx);
            psi_load<V, false>(px, c, 0, fx, lx, ((size_t)lx * ny + j) * nz + k0, c.ph_pos, c.ph_neg, c.phw_pos, c.phw_neg, TAB_BH, TAB_AH, TAB_KH);
        }
        if constexpr ((AX & AX_Y) != 0) {
            const int Ly = fy == 2 ? c.L[2] : (fy == 3 ? c.L[3] : 0);
            psi_load<V, false>(py, c, 1, fy, ly, ((size_t)i * Ly + ly) * nz + k0, c.ph_pos, c.ph_neg, c.phw_pos, c.phw_neg, TAB_BH, TAB_AH, TAB_KH);
        }
        if constexpr ((AX & AX_Z) != 0) {
            const int Lz = fz == 4 ? c.L[4] : (fz == 5 ? c.L[5] : 0);
            psi_load<V, true>(pz, c, 2, fz, lz, ((size_t)i * ny + j) * Lz + lz, c.ph_pos, c.ph_neg, c.phw_pos, c.phw_neg, TAB_BH, TAB_AH, TAB_KH);
        }

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

        // x: psi_pos <- dEz/dx (Hy, Q-term), psi_neg <- dEy/dx (Hz, P-term)
        if constexpr ((AX & AX_X) != 0) psi_apply<V, PHYS>(px, dEz_x, dEy_x, hy, b1, hz, b2, upd);
        // y: psi_pos <- dEx/dy (Hz), psi_neg <- dEz/dy (Hx)
        if constexpr ((AX & AX_Y) != 0) psi_apply<V, PHYS>(py, dEx_y, dEz_y, hz, b2, hx, b0, upd);
        // z: psi_pos <- dEy/dz (Hx), psi_neg <- dEx/dz (Hy)
        if constexpr ((AX & AX_Z) != 0) psi_apply<V, PHYS>(pz, dEy_z, dEx_z, hx, b0, hy, b1, upd);
        if constexpr ((AX & AX_X) != 0) psi_store<V>(px);
        if constexpr ((AX & AX_Y) != 0) psi_store<V>(py);
        if constexpr ((AX & AX_Z) != 0) psi_store<V>(pz);

        stn<V, OE_NT_STORE>(Hxw + p, hx);
        stn<V, OE_NT_STORE>(Hyw + p, hy);
        stn<V, OE_NT_STORE>(Hzw + p, hz);
        ey_c = ey_n;
        ez_c = ez_n;
    }
}

// Point sources of the lane's V cells (BasicEngine::apply_excitations,
// src/fdtd/basic_engine.rs:313-395): list order, soft adds, hard sets.
struct Sources {
    const PointExc* exc;
    const uint32_t* idx;  // cell indices (kernel arguments)
    int n;
    uint64_t step;
    const uint64_t* step_dev;  // non-null: read the step here (graph replay)
};

template <int V>
__device__ __forceinline__ void apply_sources(const Sources& src, size_t p, Vec<V>& ex, Vec<V>& ey, Vec<V>& ez) {
    for (int q = 0; q < src.n; ++q) {
        const size_t idx = src.idx[q];
        if (idx < p || idx >= p + V) continue;
        const PointExc* x = src.exc + q;
        const uint64_t step = src.step_dev ? *src.step_dev : src.step;
        if (step >= x->n_samples) continue;
        const float v = x->samples[step];
        const bool soft = x->soft != 0;
        const uint32_t comp = x->comp;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            if (idx != p + m) continue;
            const float fx = soft ? ex.v[m] + v : v;
            const float fy = soft ? ey.v[m] + v : v;
            const float fz = soft ? ez.v[m] + v : v;
            ex.v[m] = comp == 0 ? fx : ex.v[m];
            ey.v[m] = comp == 1 ? fy : ey.v[m];
            ez.v[m] = comp == 2 ? fz : ez.v[m];
        }
    }
}

// MUR: first-order Mur on the z faces in mz (MurAbc, src/extensions/
// mur_abc.rs:213-324) applied by the thread holding the face's line and
// shift cells (V = 4 lanes): pre v = E_s - c E_l from the old values, post
// v += c E_s from the updated ones, then E_l = v, for Ex and Ey; rows the
// update skips (i = 0 or j = 0) get the same writes from their old values.
// Valid when no other hook acts on those cells between the E update and the
// Mur post step (the host checks it).
template <int V, int AX, bool PHYS, bool TAB, bool MUR = false>
__device__ __forceinline__ void e_body(const Grid& g, const float* Fin, float* Fout, const float* __restrict__ Cf,
                                       const MatIndex mat, const float2* tab,
                                       const CpmlArgs& c, const Tile& t, const Sources* src = nullptr,
                                       const MurZ& mz = MurZ{}) {
    static_assert(!MUR || V == 4, "the folded Mur needs float4 lanes");
    const int ny = g.ny, nz = g.nz;
    if (t.k0 >= nz || t.j >= t.jend) return;
    const size_t N = g.N;
    const size_t plane = (size_t)ny * nz;
    // old E from Fin, new E to Fout; the new H (already in Fout) is read
    const bool copy_all = Fin != Fout;
    const float* Ex = Fin;
    const float* Ey = Fin + N;
    const float* Ez = Fin + 2 * N;
    float* Exw = Fout;
    float* Eyw = Fout + N;
    float* Ezw = Fout + 2 * N;
    const float* __restrict__ Hx = Fout + 3 * N;
    const float* __restrict__ Hy = Fout + 4 * N;
    const float* __restrict__ Hz = Fout + 5 * N;
    const float* __restrict__ ca0 = Cf;
    const float* __restrict__ ca1 = Cf + N;
    const float* __restrict__ ca2 = Cf + 2 * N;
    const float* __restrict__ cb0 = Cf + 3 * N;
    const float* __restrict__ cb1 = Cf + 4 * N;
    const float* __restrict__ cb2 = Cf + 5 * N;

    const int j = t.j, k0 = t.k0;
    const bool gj = j > 0;
    const bool gk_first = k0 > 0;
    int ly = 0, lz = 0;
    int fy = -1, fz = -1;
    if constexpr ((AX & AX_Y) != 0) fy = face_of(c, 1, j, ly);
    if constexpr ((AX & AX_Z) != 0) fz = face_of(c, 2, k0, lz);
    size_t p = (size_t)t.i0 * plane + (size_t)j * nz + k0;
    Vec<V> hy_p = zerov<V>(), hz_p = zerov<V>();
    if (t.i0 > 0) {
        hy_p = ldv<V>(Hy + p - plane);
        hz_p = ldv<V>(Hz + p - plane);
    }
    // lanes of the folded Mur faces held by this thread
    const bool mur_lo = MUR && mz.on[0] && k0 == 0;
    const bool mur_hi = MUR && mz.on[1] && k0 + V == nz;
    // material-table row indices one plane ahead: the LDS row reads of a
    // plane need not wait for that plane's field loads
    MatIdx mi_c;
    if constexpr (TAB) {
        if (gj && t.i0 > 0) mi_c = coef_index<V>(mat, p);
    }

    for (int i = t.i0; i < t.i1; ++i, p += plane) {
        const bool gi = i > 0;
        const bool row = gi && gj;
        MatIdx mi_n;
        if constexpr (TAB) {
            if (gj && i + 1 < t.i1) mi_n = coef_index<V>(mat, p + plane);
        }
        const Vec<V> hx_c = ldv<V>(Hx + p), hy_c = ldv<V>(Hy + p), hz_c = ldv<V>(Hz + p);
        Vec<V> hz_j = hz_c, hx_j = hx_c;
        if (gj) {
            hz_j = ldv<V>(Hz + p - nz);
            hx_j = ldv<V>(Hx + p - nz);
        }
        // the k - 1 neighbour: the previous lane's last cell, or a load where
        // that lane holds another row (or sits in another wave)
        const float hy_sh = __shfl_up(hy_c.v[V - 1], 1, 64);
        const float hx_sh = __shfl_up(hx_c.v[V - 1], 1, 64);
        float hy_kf = 0.f, hx_kf = 0.f;
        if (gk_first) {
            hy_kf = t.kfirst ? Hy[p - 1] : hy_sh;
            hx_kf = t.kfirst ? Hx[p - 1] : hx_sh;
        }
        Vec<V> ex, ey, ez, c0, c1, c2, s0, s1, s2;
        if (row) {
            ex = ldn<V, OE_NT_OWN>(Ex + p); ey = ldn<V, OE_NT_OWN>(Ey + p); ez = ldn<V, OE_NT_OWN>(Ez + p);
            if constexpr (TAB) {
                coef_from<V>(mat, mi_c, tab, s0, s1, s2, c0, c1, c2);
            } else {
                s0 = ldn<V, OE_NT_COEF>(ca0 + p); s1 = ldn<V, OE_NT_COEF>(ca1 + p); s2 = ldn<V, OE_NT_COEF>(ca2 + p);
                c0 = ldn<V, OE_NT_COEF>(cb0 + p); c1 = ldn<V, OE_NT_COEF>(cb1 + p); c2 = ldn<V, OE_NT_COEF>(cb2 + p);
            }
        }
        // Mur pre step on the old values: v = E(shift) - c E(line)
        float mvx[2] = {0.f, 0.f}, mvy[2] = {0.f, 0.f};
        if constexpr (MUR) {
            if (mur_lo || mur_hi) {
                if (!row) {
                    ex = ldn<V, OE_NT_OWN>(Ex + p);
                    ey = ldn<V, OE_NT_OWN>(Ey + p);
                }
                if (mur_lo) {
                    mvx[0] = ex.v[1] - mz.coeff[0] * ex.v[0];
                    mvy[0] = ey.v[1] - mz.coeff[0] * ey.v[0];
                }
                if (mur_hi) {
                    mvx[1] = ex.v[V - 2] - mz.coeff[1] * ex.v[V - 1];
                    mvy[1] = ey.v[V - 2] - mz.coeff[1] * ey.v[V - 1];
                }
            }
        }

        Psi<V> px, py, pz;
        if constexpr ((AX & AX_X) != 0) {
            int lx;
            const int fx = face_of(c, 0, i, lx);
            psi_load<V, false>(px, c, 0, fx, lx, ((size_t)lx * ny + j) * nz + k0, c.pe_pos, c.pe_neg, c.pe_pos, c.pe_neg, TAB_BE, TAB_AE, TAB_KE);
        }
        if constexpr ((AX & AX_Y) != 0) {
            const int Ly = fy == 2 ? c.L[2] : (fy == 3 ? c.L[3] : 0);
            psi_load<V, false>(py, c, 1, fy, ly, ((size_t)i * Ly + ly) * nz + k0, c.pe_pos, c.pe_neg, c.pe_pos, c.pe_neg, TAB_BE, TAB_AE, TAB_KE);
        }
        if constexpr ((AX & AX_Z) != 0) {
            const int Lz = fz == 4 ? c.L[4] : (fz == 5 ? c.L[5] : 0);
            psi_load<V, true>(pz, c, 2, fz, lz, ((size_t)i * ny + j) * Lz + lz, c.pe_pos, c.pe_neg, c.pe_pos, c.pe_neg, TAB_BE, TAB_AE, TAB_KE);
        }

        float dHz_y[V], dHy_z[V], dHx_z[V], dHz_x[V], dHy_x[V], dHx_y[V];
        bool upd[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const bool gk = (e > 0) || gk_first;
            upd[e] = row && gk;
            const float hyk = (e > 0) ? hy_c.v[(e > 0) ? e - 1 : e] : hy_kf;
            const float hxk = (e > 0) ? hx_c.v[(e > 0) ? e - 1 : e] : hx_kf;
            dHz_y[e] = gj ? hz_c.v[e] - hz_j.v[e] : 0.f;
            dHy_z[e] = gk ? hy_c.v[e] - hyk : 0.f;
            dHx_z[e] = gk ? hx_c.v[e] - hxk : 0.f;
            dHz_x[e] = gi ? hz_c.v[e] - hz_p.v[e] : 0.f;
            dHy_x[e] = gi ? hy_c.v[e] - hy_p.v[e] : 0.f;
            dHx_y[e] = gj ? hx_c.v[e] - hx_j.v[e] : 0.f;
        }

        if (row) {
#pragma unroll
            for (int e = 0; e < V; ++e) {
                if (!upd[e]) continue;  // k == 0: not written
                const float cx = dHz_y[e] - dHy_z[e];
                ex.v[e] = s0.v[e] * ex.v[e] + c0.v[e] * cx;
                const float cy = dHx_z[e] - dHz_x[e];
                ey.v[e] = s1.v[e] * ey.v[e] + c1.v[e] * cy;
                const float cz = dHy_x[e] - dHx_y[e];
                ez.v[e] = s2.v[e] * ez.v[e] + c2.v[e] * cz;
            }
        }

        // x: psi_pos <- dHz/dx (Ey, Q-term), psi_neg <- dHy/dx (Ez, P-term)
        if constexpr ((AX & AX_X) != 0) psi_apply<V, PHYS>(px, dHz_x, dHy_x, ey, c1, ez, c2, upd);
        // y: psi_pos <- dHx/dy (Ez), psi_neg <- dHz/dy (Ex)
        if constexpr ((AX & AX_Y) != 0) psi_apply<V, PHYS>(py, dHx_y, dHz_y, ez, c2, ex, c0, upd);
        // z: psi_pos <- dHy/dz (Ex), psi_neg <- dHx/dz (Ey)
        if constexpr ((AX & AX_Z) != 0) psi_apply<V, PHYS>(pz, dHy_z, dHx_z, ex, c0, ey, c1, upd);
        if constexpr ((AX & AX_X) != 0) psi_store<V>(px);
        if constexpr ((AX & AX_Y) != 0) psi_store<V>(py);
        if constexpr ((AX & AX_Z) != 0) psi_store<V>(pz);

        // Mur post step and boundary write: v += c E(shift), E(line) = v
        if constexpr (MUR) {
            if (mur_lo) {
                ex.v[0] = mvx[0] + mz.coeff[0] * ex.v[1];
                ey.v[0] = mvy[0] + mz.coeff[0] * ey.v[1];
            }
            if (mur_hi) {
                ex.v[V - 1] = mvx[1] + mz.coeff[1] * ex.v[V - 2];
                ey.v[V - 1] = mvy[1] + mz.coeff[1] * ey.v[V - 2];
            }
            if (!row && (mur_lo || mur_hi)) {
                stn<V, OE_NT_STORE>(Exw + p, ex);
                stn<V, OE_NT_STORE>(Eyw + p, ey);
            }
        }

        if (row) {
            if (src) apply_sources<V>(*src, p, ex, ey, ez);
            stn<V, OE_NT_STORE>(Exw + p, ex);
            stn<V, OE_NT_STORE>(Eyw + p, ey);
            stn<V, OE_NT_STORE>(Ezw + p, ez);
        } else if (copy_all) {  // index-0 planes and rows keep their value
            stn<V, OE_NT_STORE>(Exw + p, ldn<V, OE_NT_OWN>(Ex + p));
            stn<V, OE_NT_STORE>(Eyw + p, ldn<V, OE_NT_OWN>(Ey + p));
            stn<V, OE_NT_STORE>(Ezw + p, ldn<V, OE_NT_OWN>(Ez + p));
        }
        hy_p = hy_c;
        hz_p = hz_c;
        mi_c = mi_n;
    }
}

// One grid: the tile comes from the launch's box list.  With TAB the
// work-group first copies its half's material rows into LDS (dynamic LDS of
// 24 bytes per row).
template <int V, int AX, bool PHYS, bool TAB>
__global__ __launch_bounds__(256) void k_update_h(YeeArgs a, CpmlArgs c) {
    extern __shared__ float2 lds_rows[];
    // the step's first launch advances the step counter (no H lane reads it)
    if (a.step_inc && blockIdx.x == 0 && threadIdx.x == 0) *a.step_inc += 1;
    if constexpr (TAB) {
        const float2* src = reinterpret_cast<const float2*>(a.tab_h);
        for (int q = threadIdx.x; q < 3 * a.n_mat_h; q += blockDim.x) lds_rows[q] = src[q];
        __syncthreads();
    }
    h_body<V, AX, PHYS, TAB>(a.g, a.Fin, a.Fout, a.Cf, MatIndex{a.n_mat_h == 1 ? nullptr : a.mat_h, a.mat8_h}, lds_rows, c,
                             decode<V>(a));
}

template <int V, int AX, bool PHYS, bool TAB, bool MUR = false>
__global__ __launch_bounds__(256) void k_update_e(YeeArgs a, CpmlArgs c) {
    extern __shared__ float2 lds_rows[];
    if constexpr (TAB) {
        const float2* src = reinterpret_cast<const float2*>(a.tab_e);
        for (int q = threadIdx.x; q < 3 * a.n_mat_e; q += blockDim.x) lds_rows[q] = src[q];
        __syncthreads();
    }
    const MatIndex mat_e{a.n_mat_e == 1 ? nullptr : a.mat_e, a.mat8_e};
    if constexpr (MUR) {
        // sources follow the Mur step, so they are never folded with it
        e_body<V, AX, PHYS, TAB, true>(a.g, a.Fin, a.Fout, a.Cf, mat_e, lds_rows, c, decode<V>(a), nullptr, a.mur_z);
        return;
    }
    if (a.n_exc) {
        const Sources src{a.exc, a.exc_idx, a.n_exc, a.exc_step, a.exc_step_dev};
        e_body<V, AX, PHYS, TAB>(a.g, a.Fin, a.Fout, a.Cf, mat_e, lds_rows, c, decode<V>(a), &src);
    } else {
        e_body<V, AX, PHYS, TAB>(a.g, a.Fin, a.Fout, a.Cf, mat_e, lds_rows, c, decode<V>(a));
    }
}

// Many grids in one launch (a parameter sweep): work-group -> box of some
// structure through the launch's table; each box has its own tile shape
// (bx * by * ns <= 256 lanes: ns i-chunks of by rows of bx lanes; surplus
// lanes exit).
template <int V>
__device__ __forceinline__ bool decode_sweep(const SweepArgs& a, Tile& r, int& sid) {
    int bid = blockIdx.x;
    if (a.xcd_remap) {
        const int per = gridDim.x >> 3;
        if (bid < (per << 3)) bid = (bid & 7) * per + (bid >> 3);
    }
    const SweepBox& B = a.box[a.wg_box[bid]];
    bid -= B.first;
    const int kb = bid % B.n_kb;
    const int t = bid / B.n_kb;
    const int jb = t % B.n_jb;
    const int ib = t / B.n_jb;
    const int kx = threadIdx.x % B.bx;
    const int rest = threadIdx.x / B.bx;
    const int jy = rest % B.by;
    const int sub = rest / B.by;  // i-chunk within the work-group
    r.k0 = (kb * B.bx + kx) * V;
    r.j = B.j0 + jb * B.by + jy;
    r.i0 = B.i0 + (ib * B.ns + sub) * B.ci;
    r.i1 = min(r.i0 + B.ci, B.i1);
    // surplus lanes (beyond ns chunks, or a chunk past the box) get an empty
    // row range and return before any load
    r.jend = (sub < B.ns && r.i0 < B.i1) ? B.j1 : r.j;
    r.kfirst = kx == 0 || (threadIdx.x & 63) == 0;
    r.klast = kx == B.bx - 1 || (threadIdx.x & 63) == 63;
    sid = B.sid;
    return true;
}

// A work-group belongs to one structure, so with TAB it copies that
// structure's rows into LDS (every lane reaches the barrier: surplus lanes
// only get an empty row range).
template <int V, int AX, bool PHYS, bool TAB>
__global__ __launch_bounds__(256) void k_update_h_sweep(SweepArgs a) {
    extern __shared__ float2 lds_rows[];
    if (a.step_inc && blockIdx.x == 0 && threadIdx.x == 0) *a.step_inc += 1;
    Tile t;
    int sid;
    decode_sweep<V>(a, t, sid);
    const SweepStruct& S = a.st[sid];
    if constexpr (TAB) {
        const float2* src = reinterpret_cast<const float2*>(S.tab_h);
        for (int q = threadIdx.x; q < 3 * S.n_mat_h; q += blockDim.x) lds_rows[q] = src[q];
        __syncthreads();
    }
    h_body<V, AX, PHYS, TAB>(S.g, S.F, S.F, S.Cf, MatIndex{S.n_mat_h == 1 ? nullptr : S.mat_h, S.mat8_h}, lds_rows, S.c, t);
}

template <int V, int AX, bool PHYS, bool TAB>
__global__ __launch_bounds__(256) void k_update_e_sweep(SweepArgs a) {
    extern __shared__ float2 lds_rows[];
    Tile t;
    int sid;
    decode_sweep<V>(a, t, sid);
    const SweepStruct& S = a.st[sid];
    if constexpr (TAB) {
        const float2* src = reinterpret_cast<const float2*>(S.tab_e);
        for (int q = threadIdx.x; q < 3 * S.n_mat_e; q += blockDim.x) lds_rows[q] = src[q];
        __syncthreads();
    }
    e_body<V, AX, PHYS, TAB>(S.g, S.F, S.F, S.Cf, MatIndex{S.n_mat_e == 1 ? nullptr : S.mat_e, S.mat8_e}, lds_rows, S.c, t);
}

template <int V, int AX, bool PHYS>
hipError_t launch_h_t(const YeeArgs& a, const CpmlArgs& c, hipStream_t s) {
    if (a.n_mat_h > 0)
        hipLaunchKernelGGL((k_update_h<V, AX, PHYS, true>), dim3(a.first[a.nbox]), dim3(256),
                           3 * a.n_mat_h * sizeof(float2), s, a, c);
    else
        hipLaunchKernelGGL((k_update_h<V, AX, PHYS, false>), dim3(a.first[a.nbox]), dim3(256), 0, s, a, c);
    return hipGetLastError();
}

template <int V, int AX, bool PHYS>
hipError_t launch_e_t(const YeeArgs& a, const CpmlArgs& c, hipStream_t s) {
    if constexpr (V == 4) {
        if (a.mur_z.on[0] || a.mur_z.on[1]) {
            if (a.n_mat_e > 0)
                hipLaunchKernelGGL((k_update_e<V, AX, PHYS, true, true>), dim3(a.first[a.nbox]), dim3(256),
                                   3 * a.n_mat_e * sizeof(float2), s, a, c);
            else
                hipLaunchKernelGGL((k_update_e<V, AX, PHYS, false, true>), dim3(a.first[a.nbox]), dim3(256), 0, s, a,
                                   c);
            return hipGetLastError();
        }
    } else {
        if (a.mur_z.on[0] || a.mur_z.on[1]) return hipErrorInvalidValue;  // needs float4 lanes
    }
    if (a.n_mat_e > 0)
        hipLaunchKernelGGL((k_update_e<V, AX, PHYS, true>), dim3(a.first[a.nbox]), dim3(256),
                           3 * a.n_mat_e * sizeof(float2), s, a, c);
    else
        hipLaunchKernelGGL((k_update_e<V, AX, PHYS, false>), dim3(a.first[a.nbox]), dim3(256), 0, s, a, c);
    return hipGetLastError();
}

// Instantiated variants: plain (no CPML), z faces only (interior of a CPML
// grid) and all faces (the x/y face slabs), each in compat and physical
// mode, each reading the material table or the coefficient arrays.
template <bool H, int V>
hipError_t dispatch(const YeeArgs& a, const CpmlArgs* c, int axes, hipStream_t s) {
    const CpmlArgs none{};
    if (!c || axes == 0) return H ? launch_h_t<V, 0, false>(a, none, s) : launch_e_t<V, 0, false>(a, none, s);
    if (axes == AX_Z) {
        if (c->physical) return H ? launch_h_t<V, AX_Z, true>(a, *c, s) : launch_e_t<V, AX_Z, true>(a, *c, s);
        return H ? launch_h_t<V, AX_Z, false>(a, *c, s) : launch_e_t<V, AX_Z, false>(a, *c, s);
    }
    constexpr int ALL = AX_X | AX_Y | AX_Z;
    if (c->physical) return H ? launch_h_t<V, ALL, true>(a, *c, s) : launch_e_t<V, ALL, true>(a, *c, s);
    return H ? launch_h_t<V, ALL, false>(a, *c, s) : launch_e_t<V, ALL, false>(a, *c, s);
}

// ------------------------------------------------------------ fused step
// One pass per timestep over a box: H^{n+1/2} and E^{n+1} of a block of
// FUSED_ROWS k-rows (one wave each, V = 4 cells per lane) marching along i,
// reading the old fields (Fo) and writing the new ones (Fn).  The E update
// of plane i needs the new H of
//   (i, j, k-1)   the lane to the left (cross-lane shift) or its own vector,
//   (i, j-1, k)   the wave below, through LDS; for the block's first row the
//                 extra wave 0, which recomputes the H of the row below the
//                 block from the old fields (or reads it from Fn where that
//                 row is a CPML slab row, updated before this launch),
//   (i-1, j, k)   carried from the previous plane; before the first plane
//                 of the chunk it is recomputed (or read from Fn likewise).
// The arithmetic of every cell is that of h_body / e_body, so the result is
// bit-identical to the two in-place half-steps.  The z-face H psi is read
// from c.ph_* and the new values stored to c.phw_* (a second buffer), so the
// recomputations read old psi while owners write new psi.
template <int AX, bool PHYS, bool TAB>
__global__ __launch_bounds__((FUSED_ROWS + 1) * 64) void k_fused(FusedArgs a, CpmlArgs c) {
    constexpr int V = 4;
    constexpr bool ZC = (AX & AX_Z) != 0;
    extern __shared__ float4 lds4[];
    float4* xch = lds4;  // [2 parities][FUSED_ROWS writer waves][Hx, Hz][64 lanes]
    float2* tabh = reinterpret_cast<float2*>(lds4 + 2 * FUSED_ROWS * 2 * 64);
    float2* tabe = tabh + 3 * a.n_mat_h;
    const int tid = threadIdx.x, w = tid >> 6, kx = tid & 63;
    if constexpr (TAB) {
        const float2* sh = reinterpret_cast<const float2*>(a.tab_h);
        const float2* se = reinterpret_cast<const float2*>(a.tab_e);
        for (int q = tid; q < 3 * a.n_mat_h; q += blockDim.x) tabh[q] = sh[q];
        for (int q = tid; q < 3 * a.n_mat_e; q += blockDim.x) tabe[q] = se[q];
    }
    __syncthreads();

    int bid = blockIdx.x;
    if (a.xcd_remap) {
        const int per = gridDim.x >> 3;
        if (bid < (per << 3)) bid = (bid & 7) * per + (bid >> 3);
    }
    const int jb = bid % a.n_jb;
    const int ia = a.i0 + (bid / a.n_jb) * a.ci;
    const int ie = min(ia + a.ci, a.i1);
    const int j = a.j0 + jb * FUSED_ROWS - 1 + w;  // wave 0: the row below the block
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const size_t N = a.g.N;
    const size_t plane = (size_t)ny * nz;
    const int k0 = kx * V;
    const bool lane_on = k0 < nz;
    const bool owner = w > 0 && j < a.j1;
    const bool halo_comp = w == 0 && j >= a.j0;
    const bool halo_read = w == 0 && j >= 0 && j < a.j0;
    const bool comp = lane_on && (owner || halo_comp);

    const float* __restrict__ Ex = a.Fo;
    const float* __restrict__ Ey = a.Fo + N;
    const float* __restrict__ Ez = a.Fo + 2 * N;
    const float* __restrict__ Hxo = a.Fo + 3 * N;
    const float* __restrict__ Hyo = a.Fo + 4 * N;
    const float* __restrict__ Hzo = a.Fo + 5 * N;
    float* __restrict__ Exn = a.Fn;
    float* __restrict__ Eyn = a.Fn + N;
    float* __restrict__ Ezn = a.Fn + 2 * N;
    float* __restrict__ Hxn = a.Fn + 3 * N;
    float* __restrict__ Hyn = a.Fn + 4 * N;
    float* __restrict__ Hzn = a.Fn + 5 * N;

    int lz = 0, fz = -1;
    if constexpr (ZC) fz = face_of(c, 2, k0, lz);
    const int Lz = fz == 4 ? c.L[4] : (fz == 5 ? c.L[5] : 0);
    const bool hj = j + 1 < ny;
    const bool hk_last = k0 + V < nz;
    bool upd_all[V];
#pragma unroll
    for (int e = 0; e < V; ++e) upd_all[e] = true;

    // H^{n+1/2} of (i, j, k0..k0+3) from the old fields (h_body's arithmetic);
    // ey_c/ez_c hold the old Ey/Ez of plane i, ey_n/ez_n those of plane i+1.
    auto h_step = [&](int i, size_t p, const Vec<V>& ey_c, const Vec<V>& ez_c, const Vec<V>& ey_n,
                      const Vec<V>& ez_n, Vec<V>& ex_c, Vec<V>& hx, Vec<V>& hy, Vec<V>& hz, Psi<V>& pz) {
        const bool hi = i + 1 < nx;
        ex_c = ldv<V>(Ex + p);
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
        hx = ldn<V, OE_NT_OWN>(Hxo + p);
        hy = ldn<V, OE_NT_OWN>(Hyo + p);
        hz = ldn<V, OE_NT_OWN>(Hzo + p);
        Vec<V> a0, a1, a2, b0, b1, b2;
        if constexpr (TAB) {
            coef_rows<V>(MatIndex{a.n_mat_h == 1 ? nullptr : a.mat_h, nullptr}, p, tabh, a0, a1, a2, b0, b1, b2);
        } else {
            a0 = ldn<V, OE_NT_COEF>(a.Cf + 6 * N + p); a1 = ldn<V, OE_NT_COEF>(a.Cf + 7 * N + p);
            a2 = ldn<V, OE_NT_COEF>(a.Cf + 8 * N + p); b0 = ldn<V, OE_NT_COEF>(a.Cf + 9 * N + p);
            b1 = ldn<V, OE_NT_COEF>(a.Cf + 10 * N + p); b2 = ldn<V, OE_NT_COEF>(a.Cf + 11 * N + p);
        }
        if constexpr (ZC)
            psi_load<V, true>(pz, c, 2, fz, lz, ((size_t)i * ny + j) * Lz + lz, c.ph_pos, c.ph_neg, c.phw_pos,
                              c.phw_neg, TAB_BH, TAB_AH, TAB_KH);
        float dEy_z[V], dEx_z[V];
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
            dEy_z[e] = dey_dz;
            dEx_z[e] = dex_dz;
        }
        // z: psi_pos <- dEy/dz (Hx), psi_neg <- dEx/dz (Hy)
        if constexpr (ZC) psi_apply<V, PHYS>(pz, dEy_z, dEx_z, hx, b0, hy, b1, upd_all);
    };

    const size_t p0 = (size_t)ia * plane + (size_t)(j < 0 ? 0 : j) * nz + k0;
    Vec<V> ey_c = zerov<V>(), ez_c = zerov<V>();
    if (comp) {
        ey_c = ldv<V>(Ey + p0);
        ez_c = ldv<V>(Ez + p0);
    }
    // H^{n+1/2} of plane ia - 1 (its Hy, Hz enter the x-differences of plane ia)
    Vec<V> hy_p = zerov<V>(), hz_p = zerov<V>();
    if (owner && lane_on && ia > 0) {
        const size_t pp = p0 - plane;
        if (ia - 1 < a.i0) {
            hy_p = ldv<V>(Hyn + pp);
            hz_p = ldv<V>(Hzn + pp);
        } else {
            Vec<V> ex_t, hx_t;
            Psi<V> pz_t;
            h_step(ia - 1, pp, ldv<V>(Ey + pp), ldv<V>(Ez + pp), ey_c, ez_c, ex_t, hx_t, hy_p, hz_p, pz_t);
        }
    }

    for (int i = ia; i < ie; ++i) {
        const int par = (i - ia) & 1;
        const size_t p = p0 + (size_t)(i - ia) * plane;
        Vec<V> hx = zerov<V>(), hy = zerov<V>(), hz = zerov<V>(), ex_c = zerov<V>();
        Vec<V> ey_n = ey_c, ez_n = ez_c;
        Psi<V> pz;
        // the E half's coefficient rows and psi, fetched after the barrier
        // (fetching them with the H half's loads raised the register count
        // to 218 and was slower)
        Vec<V> s0, s1, s2, c0, c1, c2;
        Psi<V> pe;
        auto e_loads = [&]() {
            if constexpr (TAB) {
                coef_rows<V>(MatIndex{a.n_mat_e == 1 ? nullptr : a.mat_e, nullptr}, p, tabe, s0, s1, s2, c0, c1, c2);
            } else {
                s0 = ldn<V, OE_NT_COEF>(a.Cf + p); s1 = ldn<V, OE_NT_COEF>(a.Cf + N + p);
                s2 = ldn<V, OE_NT_COEF>(a.Cf + 2 * N + p); c0 = ldn<V, OE_NT_COEF>(a.Cf + 3 * N + p);
                c1 = ldn<V, OE_NT_COEF>(a.Cf + 4 * N + p); c2 = ldn<V, OE_NT_COEF>(a.Cf + 5 * N + p);
            }
            if constexpr (ZC)
                psi_load<V, true>(pe, c, 2, fz, lz, ((size_t)i * ny + j) * Lz + lz, c.pe_pos, c.pe_neg, c.pe_pos,
                                  c.pe_neg, TAB_BE, TAB_AE, TAB_KE);
        };
        if (comp) {
            if (i + 1 < nx) {
                ey_n = ldv<V>(Ey + p + plane);
                ez_n = ldv<V>(Ez + p + plane);
            }
            h_step(i, p, ey_c, ez_c, ey_n, ez_n, ex_c, hx, hy, hz, pz);
            if (owner) {
                stn<V, OE_NT_STORE>(Hxn + p, hx);
                stn<V, OE_NT_STORE>(Hyn + p, hy);
                stn<V, OE_NT_STORE>(Hzn + p, hz);
                if constexpr (ZC) psi_store<V>(pz);
            }
        } else if (halo_read && lane_on) {
            hx = ldv<V>(Hxn + p);
            hz = ldv<V>(Hzn + p);
        }
        if (w < FUSED_ROWS) {
            xch[((par * FUSED_ROWS + w) * 2 + 0) * 64 + kx] = make_float4(hx.v[0], hx.v[1], hx.v[2], hx.v[3]);
            xch[((par * FUSED_ROWS + w) * 2 + 1) * 64 + kx] = make_float4(hz.v[0], hz.v[1], hz.v[2], hz.v[3]);
        }
        const float hy_kf = __shfl_up(hy.v[V - 1], 1, 64);
        const float hx_kf = __shfl_up(hx.v[V - 1], 1, 64);
        __syncthreads();

        if (owner && lane_on) {
            // E^{n+1} (e_body's arithmetic)
            const bool gi = i > 0, gj = j > 0, gk_first = k0 > 0;
            const bool row = gi && gj;
            Vec<V> hx_j = hx, hz_j = hz;
            if (gj) {
                const float4 tx = xch[((par * FUSED_ROWS + (w - 1)) * 2 + 0) * 64 + kx];
                const float4 tz = xch[((par * FUSED_ROWS + (w - 1)) * 2 + 1) * 64 + kx];
                hx_j.v[0] = tx.x; hx_j.v[1] = tx.y; hx_j.v[2] = tx.z; hx_j.v[3] = tx.w;
                hz_j.v[0] = tz.x; hz_j.v[1] = tz.y; hz_j.v[2] = tz.z; hz_j.v[3] = tz.w;
            }
            e_loads();
            Vec<V> ex = ex_c, ey = ey_c, ez = ez_c;
            float dHy_z[V], dHx_z[V];
            bool upd[V];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const bool gk = (e > 0) || gk_first;
                upd[e] = row && gk;
                const float hyk = (e > 0) ? hy.v[(e > 0) ? e - 1 : e] : hy_kf;
                const float hxk = (e > 0) ? hx.v[(e > 0) ? e - 1 : e] : hx_kf;
                const float dHz_y = gj ? hz.v[e] - hz_j.v[e] : 0.f;
                dHy_z[e] = gk ? hy.v[e] - hyk : 0.f;
                dHx_z[e] = gk ? hx.v[e] - hxk : 0.f;
                const float dHz_x = gi ? hz.v[e] - hz_p.v[e] : 0.f;
                const float dHy_x = gi ? hy.v[e] - hy_p.v[e] : 0.f;
                const float dHx_y = gj ? hx.v[e] - hx_j.v[e] : 0.f;
                if (upd[e]) {
                    const float cx = dHz_y - dHy_z[e];
                    ex.v[e] = s0.v[e] * ex.v[e] + c0.v[e] * cx;
                    const float cy = dHx_z[e] - dHz_x;
                    ey.v[e] = s1.v[e] * ey.v[e] + c1.v[e] * cy;
                    const float cz = dHy_x - dHx_y;
                    ez.v[e] = s2.v[e] * ez.v[e] + c2.v[e] * cz;
                }
            }
            // z: psi_pos <- dHy/dz (Ex), psi_neg <- dHx/dz (Ey)
            if constexpr (ZC) {
                psi_apply<V, PHYS>(pe, dHy_z, dHx_z, ex, c0, ey, c1, upd);
                psi_store<V>(pe);
            }
            stn<V, OE_NT_STORE>(Exn + p, ex);
            stn<V, OE_NT_STORE>(Eyn + p, ey);
            stn<V, OE_NT_STORE>(Ezn + p, ez);
            hy_p = hy;
            hz_p = hz;
        }
        ey_c = ey_n;
        ez_c = ez_n;
    }
}

// The fused step without a barrier: FUSED_ROWS waves per work-group, each
// the owner of one k-row j, marching along i.  Its E update needs the new Hx
// and Hz of row j - 1; the wave recomputes them from the old fields instead
// of taking them from the wave below (that row's owner computes the same
// values with the same arithmetic), so no wave waits for another.  The
// recomputation reads the old z-face H psi (c.ph_*, the owners write
// c.phw_*) and stores nothing.  Rows and planes just outside the box (CPML
// slab rows and planes, whose new H the slab launch wrote to Fn before this
// one) are read from Fn.  The E update's row indices are loaded one plane
// ahead.
template <int AX, bool PHYS, bool TAB>
__global__ __launch_bounds__(FUSED_ROWS * 64) void k_fused_rows(FusedArgs a, CpmlArgs c) {
    constexpr int V = 4;
    constexpr bool ZC = (AX & AX_Z) != 0;
    extern __shared__ float2 ftab[];
    float2* tabh = ftab;
    float2* tabe = ftab + 3 * a.n_mat_h;
    const int tid = threadIdx.x, w = tid >> 6, kx = tid & 63;
    if constexpr (TAB) {
        const float2* sh = reinterpret_cast<const float2*>(a.tab_h);
        const float2* se = reinterpret_cast<const float2*>(a.tab_e);
        for (int q = tid; q < 3 * a.n_mat_h; q += blockDim.x) tabh[q] = sh[q];
        for (int q = tid; q < 3 * a.n_mat_e; q += blockDim.x) tabe[q] = se[q];
        __syncthreads();
    }
    int bid = blockIdx.x;
    if (a.xcd_remap) {
        const int per = gridDim.x >> 3;
        if (bid < (per << 3)) bid = (bid & 7) * per + (bid >> 3);
    }
    const int jb = bid % a.n_jb;
    const int ia = a.i0 + (bid / a.n_jb) * a.ci;
    const int ie = min(ia + a.ci, a.i1);
    const int j = a.j0 + jb * FUSED_ROWS + w;
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const int k0 = kx * V;
    // no barrier follows: a lane past the row or the box leaves (the k - 1
    // shuffles read lower lanes only)
    if (j >= a.j1 || k0 >= nz) return;
    const size_t N = a.g.N;
    const size_t plane = (size_t)ny * nz;
    const float* __restrict__ Ex = a.Fo;
    const float* __restrict__ Ey = a.Fo + N;
    const float* __restrict__ Ez = a.Fo + 2 * N;
    const float* __restrict__ Hxo = a.Fo + 3 * N;
    const float* __restrict__ Hyo = a.Fo + 4 * N;
    const float* __restrict__ Hzo = a.Fo + 5 * N;
    float* __restrict__ Exn = a.Fn;
    float* __restrict__ Eyn = a.Fn + N;
    float* __restrict__ Ezn = a.Fn + 2 * N;
    float* __restrict__ Hxn = a.Fn + 3 * N;
    float* __restrict__ Hyn = a.Fn + 4 * N;
    float* __restrict__ Hzn = a.Fn + 5 * N;
    const MatIndex mh{a.n_mat_h == 1 ? nullptr : a.mat_h, nullptr};
    const MatIndex me{a.n_mat_e == 1 ? nullptr : a.mat_e, a.mat8_e};

    int lz = 0, fz = -1;
    if constexpr (ZC) fz = face_of(c, 2, k0, lz);
    const int Lz = fz == 4 ? c.L[4] : (fz == 5 ? c.L[5] : 0);
    const bool gj = j > 0;
    const bool hk_last = k0 + V < nz;
    const bool gk_first = k0 > 0;
    const bool m_read = gj && j - 1 < a.j0;  // the row below is a slab row
    const bool m_comp = gj && !m_read;
    bool upd_all[V];
#pragma unroll
    for (int e = 0; e < V; ++e) upd_all[e] = true;

    auto h_coef = [&](size_t p, Vec<V>& a0, Vec<V>& a1, Vec<V>& a2, Vec<V>& b0, Vec<V>& b1, Vec<V>& b2) {
        if constexpr (TAB) {
            coef_rows<V>(mh, p, tabh, a0, a1, a2, b0, b1, b2);
        } else {
            a0 = ldn<V, OE_NT_COEF>(a.Cf + 6 * N + p); a1 = ldn<V, OE_NT_COEF>(a.Cf + 7 * N + p);
            a2 = ldn<V, OE_NT_COEF>(a.Cf + 8 * N + p); b0 = ldn<V, OE_NT_COEF>(a.Cf + 9 * N + p);
            b1 = ldn<V, OE_NT_COEF>(a.Cf + 10 * N + p); b2 = ldn<V, OE_NT_COEF>(a.Cf + 11 * N + p);
        }
    };
    // H^{n+1/2} of (i, j, k0..k0+3) from the old fields (h_body's
    // arithmetic); ey_c/ez_c hold the old Ey/Ez of plane i, ey_n/ez_n those
    // of plane i+1; ex_c returns the old Ex of (i, j).
    auto h_full = [&](int i, size_t p, const Vec<V>& ey_c, const Vec<V>& ez_c, const Vec<V>& ey_n,
                      const Vec<V>& ez_n, Vec<V>& ex_c, Vec<V>& hx, Vec<V>& hy, Vec<V>& hz, Psi<V>& pz) {
        const bool hi = i + 1 < nx;
        const bool hj = j + 1 < ny;
        ex_c = ldv<V>(Ex + p);
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
        hx = ldn<V, OE_NT_OWN>(Hxo + p);
        hy = ldn<V, OE_NT_OWN>(Hyo + p);
        hz = ldn<V, OE_NT_OWN>(Hzo + p);
        Vec<V> a0, a1, a2, b0, b1, b2;
        h_coef(p, a0, a1, a2, b0, b1, b2);
        if constexpr (ZC)
            psi_load<V, true>(pz, c, 2, fz, lz, ((size_t)i * ny + j) * Lz + lz, c.ph_pos, c.ph_neg, c.phw_pos,
                              c.phw_neg, TAB_BH, TAB_AH, TAB_KH);
        float dEy_z[V], dEx_z[V];
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
            dEy_z[e] = dey_dz;
            dEx_z[e] = dex_dz;
        }
        // z: psi_pos <- dEy/dz (Hx), psi_neg <- dEx/dz (Hy)
        if constexpr (ZC) psi_apply<V, PHYS>(pz, dEy_z, dEx_z, hx, b0, hy, b1, upd_all);
    };
    // The new Hx, Hz of row j - 1 at plane i (the same arithmetic as its
    // owner's h_full): eym_c/eym_n the old Ey of that row at planes i, i+1;
    // ez_j/ex_j the old Ez/Ex of row j at plane i.
    auto h_below = [&](int i, size_t pm, const Vec<V>& eym_c, const Vec<V>& eym_n, const Vec<V>& ez_j,
                       const Vec<V>& ex_j, Vec<V>& hx, Vec<V>& hz) {
        const bool hi = i + 1 < nx;
        const Vec<V> exm = ldv<V>(Ex + pm);
        const Vec<V> ezm = ldv<V>(Ez + pm);
        const float ey_kl = hk_last ? Ey[pm + V] : 0.f;
        hx = ldn<V, OE_NT_OWN>(Hxo + pm);
        hz = ldn<V, OE_NT_OWN>(Hzo + pm);
        Vec<V> a0, a1, a2, b0, b1, b2;
        h_coef(pm, a0, a1, a2, b0, b1, b2);
        Psi<V> pz;
        if constexpr (ZC)
            psi_load<V, true>(pz, c, 2, fz, lz, ((size_t)i * ny + (j - 1)) * Lz + lz, c.ph_pos, c.ph_neg,
                              c.phw_pos, c.phw_neg, TAB_BH, TAB_AH, TAB_KH);
        float dEy_z[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const bool hk = (e < V - 1) || hk_last;
            const float eyk = (e < V - 1) ? eym_c.v[(e < V - 1) ? e + 1 : e] : ey_kl;
            const float dez_dy = ez_j.v[e] - ezm.v[e];  // row j - 1 + 1 = j < ny
            const float dey_dz = hk ? eyk - eym_c.v[e] : 0.f;
            const float cx = dez_dy - dey_dz;
            hx.v[e] = a0.v[e] * hx.v[e] + b0.v[e] * cx;
            const float dey_dx = hi ? eym_n.v[e] - eym_c.v[e] : 0.f;
            const float dex_dy = ex_j.v[e] - exm.v[e];
            const float cz = dey_dx - dex_dy;
            hz.v[e] = a2.v[e] * hz.v[e] + b2.v[e] * cz;
            dEy_z[e] = dey_dz;
        }
        if constexpr (ZC) {
            // the Hx correction only (the Hy of this row is not needed)
            Vec<V> hy_unused = hx;
            psi_apply<V, PHYS>(pz, dEy_z, dEy_z, hx, b0, hy_unused, b1, upd_all);
        }
    };

    const size_t p0 = (size_t)ia * plane + (size_t)j * nz + k0;
    Vec<V> ey_c = ldv<V>(Ey + p0), ez_c = ldv<V>(Ez + p0);
    Vec<V> eym_c = zerov<V>();
    if (m_comp) eym_c = ldv<V>(Ey + p0 - nz);
    // H^{n+1/2} of plane ia - 1 (its Hy, Hz enter the x-differences of plane ia)
    Vec<V> hy_p = zerov<V>(), hz_p = zerov<V>();
    if (ia > 0) {
        const size_t pp = p0 - plane;
        if (ia - 1 < a.i0) {
            hy_p = ldv<V>(Hyn + pp);
            hz_p = ldv<V>(Hzn + pp);
        } else {
            Vec<V> ex_t, hx_t;
            Psi<V> pz_t;
            h_full(ia - 1, pp, ldv<V>(Ey + pp), ldv<V>(Ez + pp), ey_c, ez_c, ex_t, hx_t, hy_p, hz_p, pz_t);
        }
    }
    MatIdx mi_c;
    if constexpr (TAB) {
        if (gj && ia > 0) mi_c = coef_index<V>(me, p0);
    }

    for (int i = ia; i < ie; ++i) {
        const size_t p = p0 + (size_t)(i - ia) * plane;
        const bool hi = i + 1 < nx;
        Vec<V> ey_n = ey_c, ez_n = ez_c, eym_n = eym_c;
        if (hi) {
            ey_n = ldv<V>(Ey + p + plane);
            ez_n = ldv<V>(Ez + p + plane);
            if (m_comp) eym_n = ldv<V>(Ey + p + plane - nz);
        }
        MatIdx mi_n;
        if constexpr (TAB) {
            if (gj && i + 1 < ie) mi_n = coef_index<V>(me, p + plane);
        }
        Vec<V> ex_c, hx, hy, hz;
        Psi<V> pz;
        h_full(i, p, ey_c, ez_c, ey_n, ez_n, ex_c, hx, hy, hz, pz);
        stn<V, OE_NT_STORE>(Hxn + p, hx);
        stn<V, OE_NT_STORE>(Hyn + p, hy);
        stn<V, OE_NT_STORE>(Hzn + p, hz);
        if constexpr (ZC) psi_store<V>(pz);
        Vec<V> hx_j = hx, hz_j = hz;  // the new Hx, Hz of row j - 1
        if (m_comp) {
            h_below(i, p - nz, eym_c, eym_n, ez_c, ex_c, hx_j, hz_j);
        } else if (m_read) {
            hx_j = ldv<V>(Hxn + p - nz);
            hz_j = ldv<V>(Hzn + p - nz);
        }
        const float hy_kf = __shfl_up(hy.v[V - 1], 1, 64);
        const float hx_kf = __shfl_up(hx.v[V - 1], 1, 64);

        // E^{n+1} (e_body's arithmetic)
        const bool gi = i > 0;
        const bool row = gi && gj;
        Vec<V> s0, s1, s2, c0, c1, c2;
        if constexpr (TAB) {
            coef_from<V>(me, mi_c, tabe, s0, s1, s2, c0, c1, c2);
        } else {
            s0 = ldn<V, OE_NT_COEF>(a.Cf + p); s1 = ldn<V, OE_NT_COEF>(a.Cf + N + p);
            s2 = ldn<V, OE_NT_COEF>(a.Cf + 2 * N + p); c0 = ldn<V, OE_NT_COEF>(a.Cf + 3 * N + p);
            c1 = ldn<V, OE_NT_COEF>(a.Cf + 4 * N + p); c2 = ldn<V, OE_NT_COEF>(a.Cf + 5 * N + p);
        }
        Psi<V> pe;
        if constexpr (ZC)
            psi_load<V, true>(pe, c, 2, fz, lz, ((size_t)i * ny + j) * Lz + lz, c.pe_pos, c.pe_neg, c.pe_pos, c.pe_neg,
                              TAB_BE, TAB_AE, TAB_KE);
        Vec<V> ex = ex_c, ey = ey_c, ez = ez_c;
        float dHy_z[V], dHx_z[V];
        bool upd[V];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const bool gk = (e > 0) || gk_first;
            upd[e] = row && gk;
            const float hyk = (e > 0) ? hy.v[(e > 0) ? e - 1 : e] : hy_kf;
            const float hxk = (e > 0) ? hx.v[(e > 0) ? e - 1 : e] : hx_kf;
            const float dHz_y = gj ? hz.v[e] - hz_j.v[e] : 0.f;
            dHy_z[e] = gk ? hy.v[e] - hyk : 0.f;
            dHx_z[e] = gk ? hx.v[e] - hxk : 0.f;
            const float dHz_x = gi ? hz.v[e] - hz_p.v[e] : 0.f;
            const float dHy_x = gi ? hy.v[e] - hy_p.v[e] : 0.f;
            const float dHx_y = gj ? hx.v[e] - hx_j.v[e] : 0.f;
            if (upd[e]) {
                const float cx = dHz_y - dHy_z[e];
                ex.v[e] = s0.v[e] * ex.v[e] + c0.v[e] * cx;
                const float cy = dHx_z[e] - dHz_x;
                ey.v[e] = s1.v[e] * ey.v[e] + c1.v[e] * cy;
                const float cz = dHy_x - dHx_y;
                ez.v[e] = s2.v[e] * ez.v[e] + c2.v[e] * cz;
            }
        }
        // z: psi_pos <- dHy/dz (Ex), psi_neg <- dHx/dz (Ey)
        if constexpr (ZC) {
            psi_apply<V, PHYS>(pe, dHy_z, dHx_z, ex, c0, ey, c1, upd);
            psi_store<V>(pe);
        }
        stn<V, OE_NT_STORE>(Exn + p, ex);
        stn<V, OE_NT_STORE>(Eyn + p, ey);
        stn<V, OE_NT_STORE>(Ezn + p, ez);
        hy_p = hy;
        hz_p = hz;
        ey_c = ey_n;
        ez_c = ez_n;
        eym_c = eym_n;
        mi_c = mi_n;
    }
}

template <int AX, bool PHYS, bool TAB>
hipError_t launch_fused_rows_t(const FusedArgs& a, const CpmlArgs& c, hipStream_t s) {
    const size_t lds = TAB ? 3 * (size_t)(a.n_mat_h + a.n_mat_e) * sizeof(float2) : 0;
    hipLaunchKernelGGL((k_fused_rows<AX, PHYS, TAB>), dim3(a.n_jb * a.n_ib), dim3(FUSED_ROWS * 64), lds, s, a, c);
    return hipGetLastError();
}

template <int AX, bool PHYS, bool TAB>
hipError_t launch_fused_t(const FusedArgs& a, const CpmlArgs& c, hipStream_t s) {
    const size_t lds = 2 * FUSED_ROWS * 2 * 64 * sizeof(float4) +
                       (TAB ? 3 * (size_t)(a.n_mat_h + a.n_mat_e) * sizeof(float2) : 0);
    hipLaunchKernelGGL((k_fused<AX, PHYS, TAB>), dim3(a.n_jb * a.n_ib), dim3((FUSED_ROWS + 1) * 64), lds, s, a, c);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_update_h(const YeeArgs& a, const CpmlArgs* c, int axes, hipStream_t s) {
    return a.t.vec == 4 ? dispatch<true, 4>(a, c, axes, s) : dispatch<true, 1>(a, c, axes, s);
}

hipError_t launch_update_e(const YeeArgs& a, const CpmlArgs* c, int axes, hipStream_t s) {
    return a.t.vec == 4 ? dispatch<false, 4>(a, c, axes, s) : dispatch<false, 1>(a, c, axes, s);
}

bool fused_supported(const Grid& g) { return g.nz % 4 == 0 && g.nz >= 4 && g.nz <= 256; }

hipError_t launch_fused(const FusedArgs& a, const CpmlArgs* c, hipStream_t s) {
    if (a.n_jb * a.n_ib <= 0) return hipSuccess;
    const bool tab = a.n_mat_h > 0 && a.n_mat_e > 0;
    const CpmlArgs none{};
    if (a.kind == 2) {
        if (!c)
            return tab ? launch_fused_rows_t<0, false, true>(a, none, s) : launch_fused_rows_t<0, false, false>(a, none, s);
        if (c->physical)
            return tab ? launch_fused_rows_t<AX_Z, true, true>(a, *c, s) : launch_fused_rows_t<AX_Z, true, false>(a, *c, s);
        return tab ? launch_fused_rows_t<AX_Z, false, true>(a, *c, s) : launch_fused_rows_t<AX_Z, false, false>(a, *c, s);
    }
    if (!c) return tab ? launch_fused_t<0, false, true>(a, none, s) : launch_fused_t<0, false, false>(a, none, s);
    if (c->physical)
        return tab ? launch_fused_t<AX_Z, true, true>(a, *c, s) : launch_fused_t<AX_Z, true, false>(a, *c, s);
    return tab ? launch_fused_t<AX_Z, false, true>(a, *c, s) : launch_fused_t<AX_Z, false, false>(a, *c, s);
}

// Sweep launches: physical CPML only (every structure of a sweep carries the
// Roden-Gedney correction); axes 0 (no CPML), AX_Z, the x or y slab pair
// with z, or all faces; rows > 0 selects the material-table variant.
template <bool H, int V, int AX>
static hipError_t sweep_t(const SweepArgs& a, int n_wg, int rows, hipStream_t s) {
    const size_t lds = (size_t)rows * 3 * sizeof(float2);
    if (rows > 0) {
        if (H) hipLaunchKernelGGL((k_update_h_sweep<V, AX, AX != 0, true>), dim3(n_wg), dim3(256), lds, s, a);
        else hipLaunchKernelGGL((k_update_e_sweep<V, AX, AX != 0, true>), dim3(n_wg), dim3(256), lds, s, a);
    } else {
        if (H) hipLaunchKernelGGL((k_update_h_sweep<V, AX, AX != 0, false>), dim3(n_wg), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_update_e_sweep<V, AX, AX != 0, false>), dim3(n_wg), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

template <bool H, int V>
static hipError_t sweep_dispatch(const SweepArgs& a, int n_wg, int axes, int rows, hipStream_t s) {
    switch (axes) {
        case 0: return sweep_t<H, V, 0>(a, n_wg, rows, s);
        case AX_Z: return sweep_t<H, V, AX_Z>(a, n_wg, rows, s);
        case AX_X | AX_Z: return sweep_t<H, V, AX_X | AX_Z>(a, n_wg, rows, s);
        case AX_Y | AX_Z: return sweep_t<H, V, AX_Y | AX_Z>(a, n_wg, rows, s);
        default: return sweep_t<H, V, AX_X | AX_Y | AX_Z>(a, n_wg, rows, s);
    }
}

hipError_t launch_sweep(bool h, const SweepArgs& a, int n_wg, int axes, int vec, int rows, hipStream_t s) {
    if (n_wg <= 0) return hipSuccess;
    if (h) return vec == 4 ? sweep_dispatch<true, 4>(a, n_wg, axes, rows, s) : sweep_dispatch<true, 1>(a, n_wg, axes, rows, s);
    return vec == 4 ? sweep_dispatch<false, 4>(a, n_wg, axes, rows, s) : sweep_dispatch<false, 1>(a, n_wg, axes, rows, s);
}

}  // namespace oe
