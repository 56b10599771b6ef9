// oe_nf2ff.hip -- near-field to far-field transform on gfx950.
//
// The radiation integrals of matlab/AnalyzeNF2FF.m:186-197 for one face with
// in-plane axes p, q (node coordinates x_p, x_q, trapezoid weights w_p, w_q)
// and the four tangential equivalent currents B_c(ip, iq) factor per angle a
// with direction cosines kappa = (sin t cos f, sin t sin f, cos t) as
//   G[a, iq, c] = sum_ip exp(+j k kappa_p(a) x_p[ip]) w_p[ip] B_c(ip, iq)
//   S[a, c]     = exp(+j k kappa_0(a) x_0) sum_iq w_q[iq] exp(+j k kappa_q(a) x_q[iq]) G[a, iq, c].
// G depends on the angle only through kappa_p, and only through |kappa_p| up
// to a sign: with C = cos(k |kappa_p| x_p), S = sin(k |kappa_p| x_p) and
// s = sign(kappa_p), G = (C Br - s S Bi) + j (C Bi + s S Br).  So the first
// contraction runs once per DISTINCT |kappa_p| (U values) as one real GEMM
//   [C_u; S_u]_{u} (2U x n_p)  x  [Br | Bi] (n_p x 8 n_q)
// on the FP32 matrix cores (v_mfma_f32_32x32x2f32), and p is the in-plane
// axis with the fewer distinct values (|cos t| on a (t, f) grid has at most
// n_t of them).  The second contraction, the per-angle phase and the
// projection are float64.

#include "oe_internal.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#pragma clang fp contract(off)

namespace oe {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GBM = 128, GBN = 128, GBK = 16;  // M, N padding and K step of the FP32 GEMM
constexpr int KPAD = 32;                        // K padding (a multiple of both GEMMs' K steps)
constexpr int LDS_PAD = 32;  // row stride = tile + 32 words: the two half-waves read different bank halves

// C[M x N] = A[M x K] * B[K x N], row-major float32; M, N multiples of TILE,
// K a multiple of 16 (the caller pads with zeros).  256 threads: 4 waves in
// a 2 x 2 grid, each owning a (TILE/2) x (TILE/2) sub-tile as MB x MB MFMA
// 32x32 blocks (MB = TILE/64).  The next K step's tiles are loaded into
// registers while the matrix cores work on the current one.
// MFMA 32x32x2 operands: lane l supplies A[l % 32][l / 32] and
// B[l / 32][l % 32]; accumulator register v of lane l holds
// C[8 (v / 4) + 4 (l / 32) + (v % 4)][l % 32].
template <int TILE>
__global__ __launch_bounds__(256) void k_sgemm_mfma(const float* __restrict__ A, const float* __restrict__ B,
                                                   float* __restrict__ C, int M, int N, int K) {
    constexpr int MB = TILE / 64;        // MFMA blocks per wave and dimension
    constexpr int LV = TILE / 64;        // float4 loads per thread and matrix per K step
    __shared__ float As[GBK][TILE + LDS_PAD];
    __shared__ __align__(16) float Bs[GBK][TILE + LDS_PAD];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int bm = blockIdx.y * TILE, bn = blockIdx.x * TILE;
    const int wm = (wave >> 1) * (TILE / 2), wn = (wave & 1) * (TILE / 2);
    // A: thread loads row ar, columns ac .. ac + 4 LV - 1 of the 16-wide K slice
    const int ar = tid & (TILE - 1), ac = (tid / TILE) * 4 * LV;
    // B: thread loads row br, columns bc .. bc + 4 LV - 1
    const int br = tid / (TILE / (4 * LV)), bc = (tid % (TILE / (4 * LV))) * 4 * LV;
    floatx16 acc[MB][MB];
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

    floatx4 ra[LV], rb[LV];
    const floatx4* sa = reinterpret_cast<const floatx4*>(A + (size_t)(bm + ar) * K + ac);
    const floatx4* sb = reinterpret_cast<const floatx4*>(B + (size_t)br * N + bn + bc);
#pragma unroll
    for (int q = 0; q < LV; ++q) {
        ra[q] = sa[q];
        rb[q] = sb[q];
    }
    for (int k0 = 0; k0 < K; k0 += GBK) {
#pragma unroll
        for (int q = 0; q < LV; ++q) {
            As[ac + 4 * q + 0][ar] = ra[q][0];
            As[ac + 4 * q + 1][ar] = ra[q][1];
            As[ac + 4 * q + 2][ar] = ra[q][2];
            As[ac + 4 * q + 3][ar] = ra[q][3];
            *reinterpret_cast<floatx4*>(&Bs[br][bc + 4 * q]) = rb[q];
        }
        __syncthreads();
        // prefetch the next K slice; the last step re-reads its own slice
        // (in bounds) so that the loads stay unconditional and ra/rb stay in
        // registers
        if (k0 + GBK < K) {
            sa += GBK / 4;
            sb += (size_t)GBK * N / 4;
        }
#pragma unroll
        for (int q = 0; q < LV; ++q) {
            ra[q] = sa[q];
            rb[q] = sb[q];
        }
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 2) {
            const int kr = kk + (lane >> 5);
            float av[MB], bv[MB];
#pragma unroll
            for (int q = 0; q < MB; ++q) {
                av[q] = As[kr][wm + 32 * q + (lane & 31)];
                bv[q] = Bs[kr][wn + 32 * q + (lane & 31)];
            }
#pragma unroll
            for (int mi = 0; mi < MB; ++mi)
#pragma unroll
                for (int ni = 0; ni < MB; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < MB; ++mi)
#pragma unroll
        for (int ni = 0; ni < MB; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = bm + wm + mi * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int col = bn + wn + ni * 32 + (lane & 31);
                C[(size_t)row * N + col] = acc[mi][ni][v];
            }
}

// C = A B on the FP16 matrix cores with FP32-like accuracy: each operand
// arrives split into two fp16 halves, x = hi + lo (hi = fp16(x), lo =
// fp16(x - hi): 22 significant bits together), and C accumulates
// hi_a hi_b + hi_a lo_b + lo_a hi_b in FP32; the dropped lo_a lo_b term is
// below 2^-22 of |a||b|.  A is [M][K] (k fastest, halves Ah then Al), B is
// given transposed, [N][K] (Bh then Bl), scaled by a power of two whose
// inverse bsc[1] multiplies C at the end (exact).  Tiles of 128 x 128, K
// steps of 32 (K a multiple of 32), 4 waves of 2 x 2 blocks of
// v_mfma_f32_32x32x16_f16 (gfx950): lane l takes A[l % 32][8 (l / 32) + j]
// and B[8 (l / 32) + j][l % 32], j = 0..7, and holds the 32x32 accumulator
// as the FP32 MFMA does.  In LDS the tiles keep k fastest (rows of 40
// halves), so each lane reads its 8 halves with one 16-byte load, and every
// global load is a 16-byte copy.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
constexpr int HK = 32;    // K step of the split FP16 GEMM
constexpr int HROW = 40;  // LDS row stride in halves: 32 of K plus 8 of padding

__global__ __launch_bounds__(256) void k_hgemm3(const _Float16* __restrict__ Ahi, const _Float16* __restrict__ Alo,
                                               const _Float16* __restrict__ Bhi, const _Float16* __restrict__ Blo,
                                               float* __restrict__ C, int M, int N, int K,
                                               const float* __restrict__ bsc) {
    constexpr int TILE = 128, MB = 2;
    __shared__ __align__(16) _Float16 Ah[TILE][HROW];
    __shared__ __align__(16) _Float16 Al[TILE][HROW];
    __shared__ __align__(16) _Float16 Bh[TILE][HROW];
    __shared__ __align__(16) _Float16 Bl[TILE][HROW];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int bm = blockIdx.y * TILE, bn = blockIdx.x * TILE;
    const int wm = (wave >> 1) * (TILE / 2), wn = (wave & 1) * (TILE / 2);
    // every thread copies 16 halves (2 x 16 B) of one row of each of the four tiles
    const int r = tid >> 1, kq = (tid & 1) * 16;
    floatx16 acc[MB][MB];
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int b = 0; b < MB; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
    half8 ga[4], gb[4];  // [hi 0..7, hi 8..15, lo 0..7, lo 8..15] of this thread's row segment
    auto load = [&](int k0) {
        const _Float16* pa = Ahi + (size_t)(bm + r) * K + k0 + kq;
        const _Float16* pl = Alo + (size_t)(bm + r) * K + k0 + kq;
        const _Float16* pb = Bhi + (size_t)(bn + r) * K + k0 + kq;
        const _Float16* ql = Blo + (size_t)(bn + r) * K + k0 + kq;
        ga[0] = *reinterpret_cast<const half8*>(pa);
        ga[1] = *reinterpret_cast<const half8*>(pa + 8);
        ga[2] = *reinterpret_cast<const half8*>(pl);
        ga[3] = *reinterpret_cast<const half8*>(pl + 8);
        gb[0] = *reinterpret_cast<const half8*>(pb);
        gb[1] = *reinterpret_cast<const half8*>(pb + 8);
        gb[2] = *reinterpret_cast<const half8*>(ql);
        gb[3] = *reinterpret_cast<const half8*>(ql + 8);
    };
    load(0);
    for (int k0 = 0; k0 < K; k0 += HK) {
        *reinterpret_cast<half8*>(&Ah[r][kq]) = ga[0];
        *reinterpret_cast<half8*>(&Ah[r][kq + 8]) = ga[1];
        *reinterpret_cast<half8*>(&Al[r][kq]) = ga[2];
        *reinterpret_cast<half8*>(&Al[r][kq + 8]) = ga[3];
        *reinterpret_cast<half8*>(&Bh[r][kq]) = gb[0];
        *reinterpret_cast<half8*>(&Bh[r][kq + 8]) = gb[1];
        *reinterpret_cast<half8*>(&Bl[r][kq]) = gb[2];
        *reinterpret_cast<half8*>(&Bl[r][kq + 8]) = gb[3];
        __syncthreads();
        if (k0 + HK < K) load(k0 + HK);
#pragma unroll
        for (int kk = 0; kk < HK; kk += 16) {
            const int kc = kk + 8 * (lane >> 5);
            half8 ah[MB], al[MB], bh[MB], bl[MB];
#pragma unroll
            for (int q = 0; q < MB; ++q) {
                ah[q] = *reinterpret_cast<const half8*>(&Ah[wm + 32 * q + (lane & 31)][kc]);
                al[q] = *reinterpret_cast<const half8*>(&Al[wm + 32 * q + (lane & 31)][kc]);
                bh[q] = *reinterpret_cast<const half8*>(&Bh[wn + 32 * q + (lane & 31)][kc]);
                bl[q] = *reinterpret_cast<const half8*>(&Bl[wn + 32 * q + (lane & 31)][kc]);
            }
#pragma unroll
            for (int mi = 0; mi < MB; ++mi)
#pragma unroll
                for (int ni = 0; ni < MB; ++ni) {
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    const float cs = bsc[1];
#pragma unroll
    for (int mi = 0; mi < MB; ++mi)
#pragma unroll
        for (int ni = 0; ni < MB; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = bm + wm + mi * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int col = bn + wn + ni * 32 + (lane & 31);
                C[(size_t)row * N + col] = acc[mi][ni][v] * cs;
            }
}

// B [K][N] float -> Bhi, Blo [N][K] halves of B s, s = 2^(14 - e) for the
// largest entry m = f 2^e (0.5 <= f < 1) of the pack's block maxima
// pmax[0..npmax) (s = 1 when m is 0 or not finite), so every scaled entry is
// below 2^14.  Entries outside the packed p < np, n < nv (padding, which may
// hold an earlier face's values) are written as zeros.  64 x 64 tiles through
// LDS: reads along n, writes along k.  Work-group 0 stores (s, 1 / s) to bsc.
__global__ __launch_bounds__(256) void k_bsplit(const float* __restrict__ B, int K, int N, int np, int nv,
                                               const float* __restrict__ pmax, int npmax,
                                               _Float16* __restrict__ Bhi, _Float16* __restrict__ Blo,
                                               float* __restrict__ bsc) {
    __shared__ float red[256];
    __shared__ float t[64][65];
    const int tid = threadIdx.x;
    float m = 0.f;
    for (int q = tid; q < npmax; q += 256) m = fmaxf(m, pmax[q]);
    red[tid] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmaxf(red[tid], red[tid + o]);
        __syncthreads();
    }
    float sc = 1.f, inv = 1.f;
    if (red[0] > 0.f && red[0] < 3.0e38f) {
        int e = 0;
        (void)frexpf(red[0], &e);
        sc = ldexpf(1.f, 14 - e);
        inv = ldexpf(1.f, e - 14);
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) {
        bsc[0] = sc;
        bsc[1] = inv;
    }
    const int k0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int tx = tid & 63, ty = tid >> 6;
    for (int y = ty; y < 64; y += 4) {
        const int k = k0 + y, n = n0 + tx;
        t[y][tx] = (k < np && n < nv) ? B[(size_t)k * N + n] * sc : 0.f;
    }
    __syncthreads();
    for (int y = ty; y < 64; y += 4) {
        const int n = n0 + y, k = k0 + tx;
        if (n < N && k < K) {
            const float x = t[tx][y];
            const _Float16 h = (_Float16)x;
            Bhi[(size_t)n * K + k] = h;
            Blo[(size_t)n * K + k] = (_Float16)(x - (float)h);
        }
    }
}

// Phase rows of k_phase as fp16 halves (hi, lo of the float64 value).  One
// thread per 8 consecutive columns of a row (ld is a multiple of 8); on a
// uniform grid (uniform, xp[i] = xp[0] + i dx) the thread takes one sincos
// and steps the other 7 columns by angle addition in float64.
__global__ __launch_bounds__(256) void k_phase16(_Float16* __restrict__ Ahi, _Float16* __restrict__ Alo, int ld,
                                                size_t total, int u0, int u1, int np, const double* __restrict__ uval,
                                                const double* __restrict__ xp, double k, int uniform, double dx) {
    const size_t groups = total / 8;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += (size_t)gridDim.x * blockDim.x) {
        const size_t t0 = 8 * g;
        const int r = (int)(t0 / ld), i0 = (int)(t0 - (size_t)r * ld);
        const int u = u0 + (r >> 1);
        const bool odd = (r & 1) != 0;
        half8 h, l;
        if (u < u1 && i0 < np) {
            const double ku = k * uval[u];
            double sn, cs, sd = 0.0, cd = 1.0;
            sincos(ku * xp[i0], &sn, &cs);
            if (uniform) sincos(ku * dx, &sd, &cd);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                double v = 0.0;
                if (i0 + q < np) {
                    if (!uniform && q > 0) sincos(ku * xp[i0 + q], &sn, &cs);
                    v = odd ? sn : cs;
                }
                h[q] = (_Float16)v;
                l[q] = (_Float16)(v - (double)h[q]);
                if (uniform) {
                    const double c2 = cs * cd - sn * sd;
                    sn = sn * cd + cs * sd;
                    cs = c2;
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) h[q] = l[q] = (_Float16)0.f;
        }
        *reinterpret_cast<half8*>(Ahi + t0) = h;
        *reinterpret_cast<half8*>(Alo + t0) = l;
    }
}

// The large contractions on the FP16 matrix cores when split16 (operands
// split into fp16 halves), the others in FP32.
bool use_split16(int M, int N, bool split16) { return split16 && (M / 128) * (N / 128) >= 512; }

hipError_t launch_gemm(const float* A, const float* B, float* C, int M, int N, int K, hipStream_t s) {
    if ((M / 128) * (N / 128) >= 512)
        hipLaunchKernelGGL(k_sgemm_mfma<128>, dim3(N / 128, M / 128), dim3(256), 0, s, A, B, C, M, N, K);
    else
        hipLaunchKernelGGL(k_sgemm_mfma<64>, dim3(N / 64, M / 64), dim3(256), 0, s, A, B, C, M, N, K);
    return hipGetLastError();
}

// Rows r of A (ld floats, total = rows * ld): u = u0 + r / 2 < u1 gives
// cos (even r) or sin (odd r) of k uval[u] xp[i] for i < np, zero elsewhere.
__global__ __launch_bounds__(256) void k_phase(float* __restrict__ A, int ld, size_t total, int u0, int u1, int np,
                                               const double* __restrict__ uval, const double* __restrict__ xp,
                                               double k) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / ld), i = (int)(t - (size_t)r * ld);
        const int u = u0 + (r >> 1);
        float v = 0.f;
        if (u < u1 && i < np) {
            double sn, cs;
            sincos(k * uval[u] * xp[i], &sn, &cs);
            v = (float)((r & 1) ? sn : cs);
        }
        A[t] = v;
    }
}

// Zero rows [np, K) of B (the padding the contraction's K tiles cover), so
// that whatever an earlier face or call left there -- a non-finite value
// from a diverged run included -- cannot reach a product: A's padding
// columns are zeros, and 0 * NaN would still be NaN.
__global__ __launch_bounds__(256) void k_nf2ff_pad_rows(float* __restrict__ B, int ldb, int np, int K) {
    const size_t n = (size_t)(K - np) * ldb;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256)
        B[(size_t)np * ldb + e] = 0.f;
}

// B = [Br | Bi] for one face: row ip, columns 4 iq + c and 4 nq + 4 iq + c
// hold w_p[ip] times the equivalent current c = (J_a1, J_a2, M_a1, M_a2) at
// node (i1, i2) (ip = i1, iq = i2, or swapped when swap = 1), with J = n x H
// and M = -n x E (matlab/AnalyzeNF2FF.m:186-191): J_a1 = -s H_a2,
// J_a2 = s H_a1, M_a1 = s E_a2, M_a2 = -s E_a1 for the outward normal
// n = s e_axis.  The block partial of the radiated power,
// 1/2 sum w_p w_q Re((E x H*) . n) (matlab/AnalyzeNF2FF.m:201-209), goes to
// part[blockIdx.x].
template <int KIND>
__global__ __launch_bounds__(256) void k_nf2ff_pack(const void* __restrict__ src, double scale, int n1, int n2,
                                                    int swap, int sign, const double* __restrict__ wp,
                                                    const double* __restrict__ wq, float* __restrict__ B, int ldb,
                                                    double* __restrict__ part, float* __restrict__ pmax) {
    __shared__ double red[256];
    __shared__ float redm[256];
    const int tid = threadIdx.x;
    const int q = blockIdx.x * 256 + tid;
    double p = 0.0;
    float mx = 0.f;
    if (q < n1 * n2) {
        const int i1 = q / n2, i2 = q - i1 * n2;
        const int ip = swap ? i2 : i1, iq = swap ? i1 : i2, nq = swap ? n1 : n2;
        double vr[4], vi[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (KIND == 0) {
                const double* v = reinterpret_cast<const double*>(src) + 8 * (size_t)q + 2 * c;
                vr[c] = scale * v[0];
                vi[c] = scale * v[1];
            } else {
                const float* v = reinterpret_cast<const float*>(src) + 16 * (size_t)q + 4 * c;
                vr[c] = scale * (double)v[0];
                vi[c] = scale * (double)v[1];
            }
        }
        const double sg = sign, w = wp[ip];
        const double cr[4] = {-sg * vr[3], sg * vr[2], sg * vr[1], -sg * vr[0]};
        const double ci[4] = {-sg * vi[3], sg * vi[2], sg * vi[1], -sg * vi[0]};
        float* row = B + (size_t)ip * ldb;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float br = (float)(w * cr[c]), bi = (float)(w * ci[c]);
            row[4 * iq + c] = br;
            row[4 * nq + 4 * iq + c] = bi;
            mx = fmaxf(mx, fmaxf(fabsf(br), fabsf(bi)));
        }
        p = 0.5 * w * wq[iq] * sg * (vr[0] * vr[3] + vi[0] * vi[3] - vr[1] * vr[2] - vi[1] * vi[2]);
    }
    red[tid] = p;
    redm[tid] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            red[tid] += red[tid + o];
            redm[tid] = fmaxf(redm[tid], redm[tid + o]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[blockIdx.x] = red[0];
        pmax[blockIdx.x] = redm[0];
    }
}

#ifndef REDUCE_M
#define REDUCE_M 2  // 2: 0.894 ms device per frequency; 3: 0.978; 4: 0.914 (fewer waves)
#endif
struct ReduceArgs {
    const float* G;        // GEMM output of this chunk: rows 2 (u - u0) (C) and 2 (u - u0) + 1 (S)
    int ldg;
    const int* order;      // angles sorted by u
    const int2* groups;    // (first index into order, count) of up to REDUCE_M angles of one row
    int g0, g1;            // this chunk's range of groups[]
    int u0;
    const int* uidx;       // per angle
    const float* usgn;     // per angle: sign of kappa_p
    int nq;
    const double* kapq;    // per angle
    const double* kap0;    // per angle
    const double* xq;
    const double* wq;
    int uniform;           // xq[i] = xq[0] + i dxq (phases by recurrence)
    double dxq, x0, k;
    int slot[4];           // NL slot of current c
    double* NL;            // [n_ang][6] complex
};

// Second contraction, one group of angles at a time: the angles of a group
// share one G row (the same |kappa_p|), so each G load serves all of them.
// QL lanes split the q sum (lane l takes q = l, l + QL, ...; consecutive
// lanes read consecutive G entries); the QL partial sums of every angle are
// added by a fixed xor butterfly (deterministic) and the group's first lane
// adds the results, rotated by the face-position phase, into NL.  G = GC -/+
// j GS is formed in float32 (one rounding of the matrix-core output, 6e-8
// relative); the phase products are accumulated in float64 with fused
// multiply-adds.
constexpr int QL = 16;
constexpr int REDUCE_THREADS = 256;
constexpr int REDUCE_GROUPS = REDUCE_THREADS / QL;

__global__ __launch_bounds__(REDUCE_THREADS) void k_nf2ff_reduce(ReduceArgs r) {
    constexpr int M = REDUCE_M;
    const int ql = threadIdx.x & (QL - 1);
    const int gi = r.g0 + blockIdx.x * REDUCE_GROUPS + (int)(threadIdx.x / QL);
    const bool act = gi < r.g1;
    const int2 grp = r.groups[act ? gi : r.g0];
    const int n = act ? grp.y : 0;
    int am[M];
    float sgf[M];
    double pr[M], pi[M], str[M], sti[M];
    double sr[M][4], si[M][4];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        am[m] = r.order[grp.x + min(m, grp.y - 1)];
        sgf[m] = r.usgn[am[m]];
        const double kk = r.k * r.kapq[am[m]];
        pr[m] = 1.0; pi[m] = 0.0; str[m] = 1.0; sti[m] = 0.0;
        if (r.uniform && ql < r.nq) {
            sincos(kk * r.xq[ql], &pi[m], &pr[m]);
            sincos(kk * (QL * r.dxq), &sti[m], &str[m]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) sr[m][c] = si[m][c] = 0.0;
    }
    const float* gc = r.G + (size_t)(2 * (r.uidx[am[0]] - r.u0)) * r.ldg;
    const float* gs = gc + r.ldg;
    auto body = [&](int iq, bool uniform) {
        const float4 cr4 = *reinterpret_cast<const float4*>(gc + 4 * iq);
        const float4 ci4 = *reinterpret_cast<const float4*>(gc + 4 * r.nq + 4 * iq);
        const float4 sr4 = *reinterpret_cast<const float4*>(gs + 4 * iq);
        const float4 si4 = *reinterpret_cast<const float4*>(gs + 4 * r.nq + 4 * iq);
        const double w = r.wq[iq];
        const float gcr[4] = {cr4.x, cr4.y, cr4.z, cr4.w}, gci[4] = {ci4.x, ci4.y, ci4.z, ci4.w};
        const float gsr[4] = {sr4.x, sr4.y, sr4.z, sr4.w}, gsi[4] = {si4.x, si4.y, si4.z, si4.w};
#pragma unroll
        for (int m = 0; m < M; ++m) {
            double cs = pr[m], sn = pi[m];
            if (uniform) {
                const double t = pr[m] * str[m] - pi[m] * sti[m];
                pi[m] = pr[m] * sti[m] + pi[m] * str[m];
                pr[m] = t;
            } else {
                sincos(r.k * r.kapq[am[m]] * r.xq[iq], &sn, &cs);
            }
            const double wr = w * cs, wi = w * sn;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double gr = (double)(gcr[c] - sgf[m] * gsi[c]);
                const double gi = (double)(gci[c] + sgf[m] * gsr[c]);
                sr[m][c] = fma(wr, gr, fma(-wi, gi, sr[m][c]));
                si[m][c] = fma(wr, gi, fma(wi, gr, si[m][c]));
            }
        }
    };
    if (r.uniform) {
        for (int iq = ql; iq < r.nq; iq += QL) body(iq, true);
    } else {
        for (int iq = ql; iq < r.nq; iq += QL) body(iq, false);
    }
#pragma unroll
    for (int o = QL / 2; o > 0; o >>= 1)
#pragma unroll
        for (int m = 0; m < M; ++m)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                sr[m][c] += __shfl_xor(sr[m][c], o, QL);
                si[m][c] += __shfl_xor(si[m][c], o, QL);
            }
    if (ql != 0) return;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        if (m >= n) break;
        const int a = am[m];
        double sn, cs;
        sincos(r.k * r.kap0[a] * r.x0, &sn, &cs);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double* o = r.NL + ((size_t)a * 6 + r.slot[c]) * 2;
            o[0] += cs * sr[m][c] - sn * si[m][c];
            o[1] += cs * si[m][c] + sn * sr[m][c];
        }
    }
}

// Second contraction over mirror groups: the angles of one G row (one |kappa_p|)
// whose |kappa_q| also agree, e.g. (theta, pi - theta) x (phi, -phi, pi - phi,
// pi + phi) on a regular grid.  With sigma = sign kappa_p and s = sign kappa_q,
// a member's sum
//   sum_q w_q e^{j s k |kappa_q| x_q} (GC_q + j sigma GS_q)
//     = P1 + j sigma P2 + j s P3 - s sigma P4,
//   P1 = sum w cos GC, P2 = sum w cos GS, P3 = sum w sin GC, P4 = sum w sin GS,
// needs only the group's four partial sums (phase k |kappa_q| x_q of the
// group's first member; members agree to 1e-12 in |kappa_q|).  Each G row is
// read once per group, and the float64 work per q is that of one angle for
// the whole group.
//
// One wave per group: lane = 4 ql + c takes component c at q = ql, ql + MQ,
// ... (the four lanes of one ql read 16 consecutive bytes, the wave 256), with
// eight float64 accumulators, and issues the loads of four q steps before
// using them (the reduction is bound by the latency of these loads).  An xor
// butterfly over the MQ lanes of each component gives every lane its
// component's sums; lane (ql, c) then finishes component c of member ql (a
// group holds at most MQ members).
struct MirrorArgs {
    const int4* groups;   // (u, first index into members, count, 0)
    const double* gkq;    // per group: |kappa_q| of its first member
    const int* members;   // angles, grouped
    int m0, m1;           // this chunk's range of groups[]
};
constexpr int MQ = 16;
constexpr int MIRROR_THREADS = 256;
constexpr int MIRROR_GROUPS = MIRROR_THREADS / (4 * MQ);

// UNI: uniform q grid (phases by recurrence, r.uniform set)
template <bool UNI>
__global__ __launch_bounds__(MIRROR_THREADS) void k_nf2ff_reduce_mirror(ReduceArgs r, MirrorArgs m) {
    const int gi = m.m0 + blockIdx.x * MIRROR_GROUPS + (int)(threadIdx.x / (4 * MQ));
    if (gi >= m.m1) return;  // whole waves: no lane of another group waits on this one
    const int c = threadIdx.x & 3;
    const int ql = (threadIdx.x >> 2) & (MQ - 1);
    const int4 grp = m.groups[gi];
    const double kk = r.k * m.gkq[gi];
    double pr = 1.0, pi = 0.0, str = 1.0, sti = 0.0;
    if (UNI && ql < r.nq) {
        sincos(kk * r.xq[ql], &pi, &pr);
        sincos(kk * (MQ * r.dxq), &sti, &str);
    }
    // P1..P4 of component c, real and imaginary
    double p1r = 0.0, p1i = 0.0, p2r = 0.0, p2i = 0.0, p3r = 0.0, p3i = 0.0, p4r = 0.0, p4i = 0.0;
    const float* gc = r.G + (size_t)(2 * (grp.x - r.u0)) * r.ldg + c;
    const float* gs = gc + r.ldg;
    const int n4 = 4 * r.nq;
    auto step = [&](int q, float a, float b, float e, float f, double w) {
        double cs = pr, sn = pi;
        if (UNI) {
            const double t = pr * str - pi * sti;
            pi = pr * sti + pi * str;
            pr = t;
        } else {
            sincos(kk * r.xq[q], &sn, &cs);
        }
        const double wc = w * cs, ws = w * sn;
        p1r = fma(wc, (double)a, p1r);
        p1i = fma(wc, (double)b, p1i);
        p2r = fma(wc, (double)e, p2r);
        p2i = fma(wc, (double)f, p2i);
        p3r = fma(ws, (double)a, p3r);
        p3i = fma(ws, (double)b, p3i);
        p4r = fma(ws, (double)e, p4r);
        p4i = fma(ws, (double)f, p4i);
    };
    int iq = ql;
    for (; iq + 3 * MQ < r.nq; iq += 4 * MQ) {
        float va[4], vb[4], ve[4], vf[4];
        double vw[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = iq + u * MQ;
            va[u] = gc[4 * q];
            vb[u] = gc[n4 + 4 * q];
            ve[u] = gs[4 * q];
            vf[u] = gs[n4 + 4 * q];
            vw[u] = r.wq[q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) step(iq + u * MQ, va[u], vb[u], ve[u], vf[u], vw[u]);
    }
    for (; iq < r.nq; iq += MQ) step(iq, gc[4 * iq], gc[n4 + 4 * iq], gs[4 * iq], gs[n4 + 4 * iq], r.wq[iq]);
#pragma unroll
    for (int o = 4; o < 4 * MQ; o <<= 1) {
        p1r += __shfl_xor(p1r, o);
        p1i += __shfl_xor(p1i, o);
        p2r += __shfl_xor(p2r, o);
        p2i += __shfl_xor(p2i, o);
        p3r += __shfl_xor(p3r, o);
        p3i += __shfl_xor(p3i, o);
        p4r += __shfl_xor(p4r, o);
        p4i += __shfl_xor(p4i, o);
    }
    const int slot = c == 0 ? r.slot[0] : c == 1 ? r.slot[1] : c == 2 ? r.slot[2] : r.slot[3];
    if (ql < grp.z) {
        const int a = m.members[grp.y + ql];
        const double sg = (double)r.usgn[a];
        const double s = r.kapq[a] < 0.0 ? -1.0 : 1.0;
        const double ss = s * sg;
        double sn, cs;
        sincos(r.k * r.kap0[a] * r.x0, &sn, &cs);
        // P1 + j sigma P2 + j s P3 - s sigma P4
        const double tr = p1r - sg * p2i - s * p3i - ss * p4r;
        const double ti = p1i + sg * p2r + s * p3r - ss * p4i;
        double* o = r.NL + ((size_t)a * 6 + slot) * 2;
        o[0] += cs * tr - sn * ti;
        o[1] += cs * ti + sn * tr;
    }
}

// E_theta = -fac (L_phi + Z0 N_theta), E_phi = fac (L_theta - Z0 N_phi) with
// fac = jk e^{-jkr}/(4 pi r) (matlab/AnalyzeNF2FF.m:98-105) and
// U = r^2/(2 Z0)(|E_theta|^2 + |E_phi|^2) (:112): out = [4][n_ang]
// (E_theta re, im, E_phi re, im); the block maximum of U goes to umax[block].
__global__ __launch_bounds__(256) void k_nf2ff_project(const double* __restrict__ NL, const double* __restrict__ NL2,
                                                       const double* __restrict__ theta,
                                                       const double* __restrict__ phi, int n_phi, size_t n_ang,
                                                       double fr, double fi, double r, double* __restrict__ out,
                                                       double* __restrict__ umax) {
    __shared__ double red[256];
    const size_t a = (size_t)blockIdx.x * 256 + threadIdx.x;
    double u = 0.0;
    if (a < n_ang) {
        const int it = (int)(a / n_phi), ip = (int)(a - (size_t)it * n_phi);
        double st, ct, sp, cp;
        sincos(theta[it], &st, &ct);
        sincos(phi[ip], &sp, &cp);
        // N, L summed over the faces (NL2: the side stream's faces, or null)
        double v[12];
#pragma unroll
        for (int c = 0; c < 12; ++c) v[c] = NL[12 * a + c] + (NL2 ? NL2[12 * a + c] : 0.0);
        const double ntr = v[0] * (ct * cp) + v[2] * (ct * sp) - v[4] * st;
        const double nti = v[1] * (ct * cp) + v[3] * (ct * sp) - v[5] * st;
        const double npr = -v[0] * sp + v[2] * cp, npi = -v[1] * sp + v[3] * cp;
        const double ltr = v[6] * (ct * cp) + v[8] * (ct * sp) - v[10] * st;
        const double lti = v[7] * (ct * cp) + v[9] * (ct * sp) - v[11] * st;
        const double lpr = -v[6] * sp + v[8] * cp, lpi = -v[7] * sp + v[9] * cp;
        const double ar = lpr + Z0 * ntr, ai = lpi + Z0 * nti;  // E_theta = -fac * A
        const double br = ltr - Z0 * npr, bi = lti - Z0 * npi;  // E_phi = fac * B
        const double etr = -(fr * ar - fi * ai), eti = -(fr * ai + fi * ar);
        const double epr = fr * br - fi * bi, epi = fr * bi + fi * br;
        // interleaved complex: [E_theta][a][re, im], then [E_phi][a][re, im]
        out[2 * a] = etr;
        out[2 * a + 1] = eti;
        out[2 * n_ang + 2 * a] = epr;
        out[2 * n_ang + 2 * a + 1] = epi;
        u = r * r / (2.0 * Z0) * (etr * etr + eti * eti + epr * epr + epi * epi);
    }
    red[threadIdx.x] = u;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) umax[blockIdx.x] = red[0];
}

inline size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

std::vector<double> trapz_weights(const std::vector<double>& c) {
    std::vector<double> w(c.size(), 0.0);
    for (size_t i = 0; i + 1 < c.size(); ++i) {
        const double h = 0.5 * (c[i + 1] - c[i]);
        w[i] += h;
        w[i + 1] += h;
    }
    return w;
}

// x[i] == x[0] + i * d within 1e-9 d
bool uniform_spacing(const std::vector<double>& x, double* d) {
    *d = x.size() > 1 ? (x.back() - x.front()) / (double)(x.size() - 1) : 0.0;
    for (size_t i = 0; i < x.size(); ++i)
        if (std::fabs(x[i] - (x.front() + (double)i * *d)) > 1e-9 * std::fabs(*d)) return false;
    return true;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() = default;
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

#define NF_OK(expr)                              \
    do {                                         \
        hipError_t _e = (expr);                  \
        if (_e != hipSuccess) return _e;         \
    } while (0)

}  // namespace

// Angle plan and work buffers, kept between calls with the same angles.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    hipError_t alloc(size_t n) {
        if (p) { (void)hipHostFree(p); p = nullptr; }
        bytes = n;
        return n ? hipHostMalloc(&p, n, hipHostMallocDefault) : hipSuccess;
    }
};

struct Nf2ffCache {
    std::vector<double> theta, phi;
    size_t n_ang = 0;
    DevBuf kap;  // [3][n_ang] double, then theta[nt], phi[np]
    PinnedBuf host;  // far-field staging
    struct Axis {
        int U = 0;
        std::vector<int> start;  // U + 1 offsets into order
        std::vector<int> gstart; // U + 1 offsets into groups
        DevBuf groups;           // int2 per group: first index into order, angle count (<= REDUCE_M)
        DevBuf uval;             // U doubles (|kappa| of each distinct value)
        DevBuf uidx;             // n_ang ints
        DevBuf usgn;             // n_ang floats
        DevBuf order;            // n_ang ints, angles sorted by u
        std::vector<int> horder; // host copy of order
    } ax[3];
    // Mirror groups of the faces normal to each axis (k_nf2ff_reduce_mirror):
    // the angles of every row u of that face's first-contraction axis,
    // grouped by equal |kappa| along its second axis, at most MQ per group.
    struct Mirror {
        std::vector<int> mstart;  // U + 1 offsets into groups
        DevBuf groups;            // int4 (u, first index into members, count, 0)
        DevBuf gkq;               // double per group
        DevBuf members;           // n_ang ints
    } mir[3];
    DevBuf work;
    // timing events, created on first use and kept: [0] brackets the whole
    // transform, the others one GEMM launch each
    std::vector<std::unique_ptr<EventPair>> events;
    // byte offset and length of the B region last zeroed in `work`
    size_t bm_off[2] = {SIZE_MAX, SIZE_MAX}, bm_len[2] = {0, 0};
    // the side stream of the x/y faces, created on first use
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    Nf2ffCache() = default;
    Nf2ffCache(const Nf2ffCache&) = delete;
    Nf2ffCache& operator=(const Nf2ffCache&) = delete;
    ~Nf2ffCache() {
        if (side) (void)hipStreamSynchronize(side);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
    }
};

Nf2ffCache* nf2ff_cache_new() { return new Nf2ffCache(); }
void nf2ff_cache_free(Nf2ffCache* c) { delete c; }

// The faces normal to axis f contract first over the in-plane axis p with
// fewer distinct |kappa| (fewer GEMM rows), then over the other, q.
static void face_axes(const Nf2ffCache& c, int f, int* p, int* q) {
    const int a1 = (f + 1) % 3, a2 = (f + 2) % 3;
    const bool swap = c.ax[a2].U < c.ax[a1].U;
    *p = swap ? a2 : a1;
    *q = swap ? a1 : a2;
}

static hipError_t plan_angles(Nf2ffCache& c, const double* theta, int nt, const double* phi, int np, hipStream_t s) {
    if (c.n_ang && c.theta.size() == (size_t)nt && c.phi.size() == (size_t)np &&
        std::equal(c.theta.begin(), c.theta.end(), theta) && std::equal(c.phi.begin(), c.phi.end(), phi))
        return hipSuccess;
    c.n_ang = 0;
    const size_t n = (size_t)nt * np;
    std::vector<double> kap(3 * n);
    for (int it = 0; it < nt; ++it)
        for (int ip = 0; ip < np; ++ip) {
            const size_t a = (size_t)it * np + ip;
            kap[a] = std::sin(theta[it]) * std::cos(phi[ip]);
            kap[n + a] = std::sin(theta[it]) * std::sin(phi[ip]);
            kap[2 * n + a] = std::cos(theta[it]);
        }
    kap.insert(kap.end(), theta, theta + nt);
    kap.insert(kap.end(), phi, phi + np);
    NF_OK(c.kap.alloc(kap.size() * sizeof(double)));
    NF_OK(hipMemcpyAsync(c.kap.p, kap.data(), kap.size() * 8, hipMemcpyHostToDevice, s));
    std::vector<int> order(n), uidx(n);
    std::vector<float> sgn(n);
    std::vector<double> uval;
    for (int x = 0; x < 3; ++x) {
        const double* kx = kap.data() + x * n;
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int p, int q) {
            const double a = std::fabs(kx[p]), b = std::fabs(kx[q]);
            return a < b || (a == b && p < q);
        });
        uval.clear();
        auto& A = c.ax[x];
        A.start.clear();
        for (size_t i = 0; i < n; ++i) {
            const double v = std::fabs(kx[order[i]]);
            // values within 1e-12 of a group's first one share its row: the
            // phase error k * 1e-12 * |x| is far below float32 resolution
            if (uval.empty() || v - uval.back() > 1e-12) {
                uval.push_back(v);
                A.start.push_back((int)i);
            }
            uidx[order[i]] = (int)uval.size() - 1;
            sgn[order[i]] = kx[order[i]] < 0.0 ? -1.f : 1.f;
        }
        A.start.push_back((int)n);
        A.U = (int)uval.size();
        // groups of up to REDUCE_M angles of one row: the reduction applies
        // every G load of the row to all of a group's angles
        std::vector<int2> groups;
        A.gstart.assign(1, 0);
        for (int u = 0; u < A.U; ++u) {
            for (int i = A.start[u]; i < A.start[u + 1]; i += REDUCE_M)
                groups.push_back(make_int2(i, std::min(REDUCE_M, A.start[u + 1] - i)));
            A.gstart.push_back((int)groups.size());
        }
        NF_OK(A.groups.alloc(groups.size() * sizeof(int2)));
        NF_OK(hipMemcpyAsync(A.groups.p, groups.data(), groups.size() * sizeof(int2), hipMemcpyHostToDevice, s));
        NF_OK(A.uval.alloc(uval.size() * 8));
        NF_OK(A.uidx.alloc(n * 4));
        NF_OK(A.usgn.alloc(n * 4));
        NF_OK(A.order.alloc(n * 4));
        NF_OK(hipMemcpyAsync(A.uval.p, uval.data(), uval.size() * 8, hipMemcpyHostToDevice, s));
        NF_OK(hipMemcpyAsync(A.uidx.p, uidx.data(), n * 4, hipMemcpyHostToDevice, s));
        NF_OK(hipMemcpyAsync(A.usgn.p, sgn.data(), n * 4, hipMemcpyHostToDevice, s));
        NF_OK(hipMemcpyAsync(A.order.p, order.data(), n * 4, hipMemcpyHostToDevice, s));
        A.horder = order;
        NF_OK(hipStreamSynchronize(s));  // host vectors are reused for the next axis
    }
    for (int f = 0; f < 3; ++f) {
        int p, q;
        face_axes(c, f, &p, &q);
        const auto& A = c.ax[p];
        const double* kq = kap.data() + (size_t)q * n;
        std::vector<int> members(n);
        std::vector<int4> groups;
        std::vector<double> gkq;
        auto& M = c.mir[f];
        M.mstart.assign(1, 0);
        for (int u = 0; u < A.U; ++u) {
            const int b = A.start[u], e = A.start[u + 1];
            for (int i = b; i < e; ++i) members[i] = A.horder[i];
            std::sort(members.begin() + b, members.begin() + e, [&](int x, int y) {
                const double a1 = std::fabs(kq[x]), a2 = std::fabs(kq[y]);
                return a1 < a2 || (a1 == a2 && x < y);
            });
            for (int i = b; i < e;) {
                const double v0 = std::fabs(kq[members[i]]);
                int j = i + 1;
                // at most MQ members: one finishing pass per lane (the poles
                // alone would otherwise make one group of every phi)
                while (j < e && j - i < MQ && std::fabs(kq[members[j]]) - v0 <= 1e-12) ++j;
                groups.push_back(make_int4(u, i, j - i, 0));
                gkq.push_back(v0);
                i = j;
            }
            M.mstart.push_back((int)groups.size());
        }
        NF_OK(M.groups.alloc(std::max<size_t>(groups.size(), 1) * sizeof(int4)));
        NF_OK(M.gkq.alloc(std::max<size_t>(gkq.size(), 1) * sizeof(double)));
        NF_OK(M.members.alloc(n * 4));
        NF_OK(hipMemcpyAsync(M.groups.p, groups.data(), groups.size() * sizeof(int4), hipMemcpyHostToDevice, s));
        NF_OK(hipMemcpyAsync(M.gkq.p, gkq.data(), gkq.size() * sizeof(double), hipMemcpyHostToDevice, s));
        NF_OK(hipMemcpyAsync(M.members.p, members.data(), n * 4, hipMemcpyHostToDevice, s));
        NF_OK(hipStreamSynchronize(s));
    }
    c.theta.assign(theta, theta + nt);
    c.phi.assign(phi, phi + np);
    c.n_ang = n;
    return hipSuccess;
}

// G stays within 64 Mi floats per GEMM launch; larger U is split into chunks
// of u, each with the angles whose u falls in it.
hipError_t nf2ff_compute(const std::vector<Nf2ffFace>& faces, double freq, const double* theta, int n_theta,
                         const double* phi, int n_phi, double r, const double center[3], hipStream_t s,
                         Nf2ffCache* cache, Nf2ffOut* out) {
    if (faces.empty() || n_theta <= 0 || n_phi <= 0 || !theta || !phi || !out || !cache || !(r > 0) || !(freq > 0))
        return hipErrorInvalidValue;
    const size_t n_ang = (size_t)n_theta * n_phi;
    if (n_ang > (size_t)INT32_MAX / 16) return hipErrorInvalidValue;
    for (const auto& f : faces) {
        const size_t n1 = f.c1.size(), n2 = f.c2.size();
        if (n1 < 1 || n2 < 1 || !f.src || f.axis < 0 || f.axis > 2 || (f.src_kind != 0 && f.src_kind != 1))
            return hipErrorInvalidValue;
        if (n1 * n2 > (size_t)INT32_MAX / 16) return hipErrorInvalidValue;
    }
    NF_OK(hipStreamSynchronize(s));
    NF_OK(plan_angles(*cache, theta, n_theta, phi, n_phi, s));
    Nf2ffCache& C = *cache;
    const double k = 2.0 * M_PI * freq / C0;
    const double* d_kap = C.kap.as<double>();

    // per face: first-contraction axis p (fewer distinct |kappa|), sizes
    struct Plan {
        int p, q, np, nq, K, N, swap;
        std::vector<double> xp, xq;
    };
    std::vector<Plan> plans;
    size_t coord_len = 0, blocks = 0;
    for (const auto& f : faces) {
        const int a1 = (f.axis + 1) % 3, a2 = (f.axis + 2) % 3;
        Plan pl{};
        pl.swap = C.ax[a2].U < C.ax[a1].U;
        pl.p = pl.swap ? a2 : a1;
        pl.q = pl.swap ? a1 : a2;
        const auto& cp = pl.swap ? f.c2 : f.c1;
        const auto& cq = pl.swap ? f.c1 : f.c2;
        pl.np = (int)cp.size();
        pl.nq = (int)cq.size();
        pl.K = (int)round_up(pl.np, KPAD);
        pl.N = (int)round_up(8 * (size_t)pl.nq, GBN);
        for (double v : cp) pl.xp.push_back(v - center[pl.p]);
        for (double v : cq) pl.xq.push_back(v - center[pl.q]);
        coord_len += 2 * ((size_t)pl.np + pl.nq);
        blocks += (cp.size() * cq.size() + 255) / 256;
        plans.push_back(std::move(pl));
    }
    // Faces run in up to two lanes: the z faces (the large contractions) on
    // the caller's stream, the x and y faces on a side stream with buffers
    // and an N/L accumulator of their own, so that their small,
    // latency-bound launches overlap the large ones (OE_NF2FF_STREAMS=1: all
    // on the caller's stream).
    int n_streams = 2;
    if (const char* v = getenv("OE_NF2FF_STREAMS")) n_streams = atoi(v) == 1 ? 1 : 2;
    std::vector<int> lane_of(faces.size(), 0);
    bool in_lane[2] = {false, false};
    for (size_t fi = 0; fi < faces.size(); ++fi) {
        lane_of[fi] = (n_streams == 2 && faces[fi].axis != 2) ? 1 : 0;
        in_lane[lane_of[fi]] = true;
    }
    if (!in_lane[0] || !in_lane[1]) std::fill(lane_of.begin(), lane_of.end(), 0);
    const int L = (in_lane[0] && in_lane[1]) ? 2 : 1;
    struct LaneBuf {
        size_t Kmax = KPAD, Nmax = GBN, rows = GBM, off = 0, bt_off = 0;
        float *Bm = nullptr, *Am = nullptr, *G = nullptr;
        _Float16* bt = nullptr;
        double* nl = nullptr;
    } lb[2];
    const size_t budget = (size_t)64 << 20;
    for (int l = 0; l < L; ++l) {
        int Umax = 0;
        for (size_t fi = 0; fi < faces.size(); ++fi) {
            if (lane_of[fi] != l) continue;
            lb[l].Kmax = std::max(lb[l].Kmax, (size_t)plans[fi].K);
            lb[l].Nmax = std::max(lb[l].Nmax, (size_t)plans[fi].N);
            Umax = std::max(Umax, C.ax[plans[fi].p].U);
        }
        const size_t r0 = std::max<size_t>(GBM, budget / lb[l].Nmax / GBM * GBM);
        lb[l].rows = std::min(r0, round_up(2 * (size_t)Umax, GBM));
    }

    const size_t pblocks = (n_ang + 255) / 256;
    const size_t n_dbl = 12 * n_ang * L + 4 * n_ang + pblocks + coord_len + blocks;
    size_t off = round_up(n_dbl * sizeof(double), 256);
    for (int l = 0; l < L; ++l) {
        lb[l].off = off;
        off = round_up(off + (lb[l].Kmax * lb[l].Nmax + lb[l].rows * lb[l].Kmax + lb[l].rows * lb[l].Nmax) *
                                 sizeof(float), 256);
        lb[l].bt_off = off;
        off = round_up(off + lb[l].Kmax * lb[l].Nmax * sizeof(float), 256);
    }
    const size_t pmax_off = off;
    off = round_up(off + blocks * sizeof(float), 256);
    const size_t bsc_off = off;
    const size_t bytes = off + 2 * faces.size() * sizeof(float);
    if (C.work.bytes < bytes) {
        NF_OK(C.work.alloc(bytes));
        C.bm_off[0] = C.bm_off[1] = SIZE_MAX;
    }
    char* wk = reinterpret_cast<char*>(C.work.p);
    double* d_nl = C.work.as<double>();
    double* d_ff = d_nl + 12 * n_ang * L;
    double* d_umax = d_ff + 4 * n_ang;
    double* d_coord = d_umax + pblocks;
    double* d_part = d_coord + coord_len;
    // B is zeroed only when its region moves or grows.  Its padding rows
    // (p >= np) are cleared per face after the pack (k_nf2ff_pad_rows), and
    // A's padding columns are exact zeros (k_phase), so they add signed zeros
    // to every product; B's padding columns only reach G columns the
    // reduction never reads.
    for (int l = 0; l < L; ++l) {
        LaneBuf& b = lb[l];
        b.Bm = reinterpret_cast<float*>(wk + b.off);
        b.Am = b.Bm + b.Kmax * b.Nmax;
        b.G = b.Am + b.rows * b.Kmax;
        // split fp16 halves of B (transposed, [N][K] each)
        b.bt = reinterpret_cast<_Float16*>(wk + b.bt_off);
        b.nl = d_nl + 12 * n_ang * l;
        const size_t len = b.Kmax * b.Nmax * sizeof(float);
        if (b.off != C.bm_off[l] || len > C.bm_len[l]) {
            NF_OK(hipMemsetAsync(b.Bm, 0, len, s));
            C.bm_off[l] = b.off;
            C.bm_len[l] = len;
        }
    }
    float* d_pmax = reinterpret_cast<float*>(wk + pmax_off);  // the pack's block maxima of |B|, per face as d_part
    // per face the scale of B and its inverse
    float* d_bsc = reinterpret_cast<float*>(wk + bsc_off);
    if (L == 2 && !C.side) {
        NF_OK(hipStreamCreateWithFlags(&C.side, hipStreamNonBlocking));
        NF_OK(hipEventCreateWithFlags(&C.ev_fork, hipEventDisableTiming));
        NF_OK(hipEventCreateWithFlags(&C.ev_join, hipEventDisableTiming));
    }
    // the large contractions on the FP16 matrix cores (operands split in two
    // halves, k_hgemm3) unless OE_NF2FF_SPLIT16=0
    bool split16 = true;
    if (const char* v = getenv("OE_NF2FF_SPLIT16")) split16 = atoi(v) != 0;
    // the second contraction by mirror groups where they hold two or more
    // angles on average, unless OE_NF2FF_MIRROR=0
    bool mirror = true;
    if (const char* v = getenv("OE_NF2FF_MIRROR")) mirror = atoi(v) != 0;

    std::vector<double> coords;
    coords.reserve(coord_len);
    for (const auto& pl : plans) {
        coords.insert(coords.end(), pl.xp.begin(), pl.xp.end());
        const auto wp = trapz_weights(pl.xp);
        coords.insert(coords.end(), wp.begin(), wp.end());
        coords.insert(coords.end(), pl.xq.begin(), pl.xq.end());
        const auto wq = trapz_weights(pl.xq);
        coords.insert(coords.end(), wq.begin(), wq.end());
    }
    NF_OK(hipMemcpyAsync(d_coord, coords.data(), coords.size() * 8, hipMemcpyHostToDevice, s));
    NF_OK(hipMemsetAsync(d_nl, 0, 12 * n_ang * L * sizeof(double), s));

    size_t n_ev = 0;
    auto next_events = [&](EventPair** out) -> hipError_t {
        if (n_ev == C.events.size()) {
            std::unique_ptr<EventPair> p(new EventPair());
            NF_OK(hipEventCreate(&p->a));
            NF_OK(hipEventCreate(&p->b));
            C.events.push_back(std::move(p));
        }
        *out = C.events[n_ev++].get();
        return hipSuccess;
    };
    EventPair* all = nullptr;
    NF_OK(next_events(&all));
    std::vector<EventPair*> gemm_ev;
    std::vector<bool> gemm_split;  // launch ran the split FP16 kernel
    double flop_split = 0.0;
    NF_OK(hipEventRecord(all->a, s));
    if (L == 2) {
        NF_OK(hipEventRecord(C.ev_fork, s));
        NF_OK(hipStreamWaitEvent(C.side, C.ev_fork, 0));
    }
    double flop = 0.0;
    // per face, its offsets into the coordinates and the pack's block sums
    std::vector<size_t> co_of(faces.size()), bo_of(faces.size());
    for (size_t fi = 0, co = 0, bo = 0; fi < faces.size(); ++fi) {
        co_of[fi] = co;
        bo_of[fi] = bo;
        co += 2 * ((size_t)plans[fi].np + plans[fi].nq);
        bo += (faces[fi].c1.size() * faces[fi].c2.size() + 255) / 256;
    }
    // the caller's stream first: its long contractions then start while
    // the host still issues the side stream's faces
    std::vector<size_t> order;
    for (int l = 0; l < L; ++l)
        for (size_t fi = 0; fi < faces.size(); ++fi)
            if (lane_of[fi] == l) order.push_back(fi);
    for (size_t fi : order) {
        const Nf2ffFace& f = faces[fi];
        // this face's lane: its stream and buffers
        const int ln = lane_of[fi];
        const hipStream_t fs = ln ? C.side : s;
        float* Bm = lb[ln].Bm;
        float* Am = lb[ln].Am;
        float* G = lb[ln].G;
        _Float16* d_bt = lb[ln].bt;
        const size_t rows = lb[ln].rows;
        const Plan& pl = plans[fi];
        const auto& A = C.ax[pl.p];
        const int n1 = (int)f.c1.size(), n2 = (int)f.c2.size();
        const int a1 = (f.axis + 1) % 3, a2 = (f.axis + 2) % 3;
        const double* xp = d_coord + co_of[fi];
        const double* wp = xp + pl.np;
        const double* xq = wp + pl.np;
        const double* wq = xq + pl.nq;
        const size_t bo = bo_of[fi];
        const int nb = (n1 * n2 + 255) / 256;
        if (f.src_kind == 0)
            hipLaunchKernelGGL(k_nf2ff_pack<0>, dim3(nb), dim3(256), 0, fs, f.src, f.scale, n1, n2, pl.swap, f.sign,
                               wp, wq, Bm, pl.N, d_part + bo, d_pmax + bo);
        else
            hipLaunchKernelGGL(k_nf2ff_pack<1>, dim3(nb), dim3(256), 0, fs, f.src, f.scale, n1, n2, pl.swap, f.sign,
                               wp, wq, Bm, pl.N, d_part + bo, d_pmax + bo);
        NF_OK(hipGetLastError());
        if (pl.K > pl.np) {
            const size_t npad = (size_t)(pl.K - pl.np) * pl.N;
            hipLaunchKernelGGL(k_nf2ff_pad_rows, dim3((unsigned)std::min<size_t>((npad + 255) / 256, 1024)), dim3(256),
                               0, fs, Bm, pl.N, pl.np, pl.K);
            NF_OK(hipGetLastError());
        }
        const float* face_pmax = d_pmax + bo;
        const int uchunk = (int)(rows / 2);
        bool face16 = false;  // some contraction of this face runs on the FP16 matrix cores
        for (int u0 = 0; u0 < A.U; u0 += uchunk) {
            const int mp = (int)round_up(2 * (size_t)(std::min(A.U, u0 + uchunk) - u0), GBM);
            face16 = face16 || use_split16(mp, pl.N, split16);
        }
        _Float16* Bhi = d_bt;
        _Float16* Blo = d_bt + (size_t)pl.N * pl.K;
        float* bsc = d_bsc + 2 * fi;
        if (face16) {
            hipLaunchKernelGGL(k_bsplit, dim3(pl.N / 64 + (pl.N % 64 != 0), (pl.K + 63) / 64), dim3(256), 0, fs, Bm, pl.K,
                               pl.N, pl.np, 8 * pl.nq, face_pmax, nb, Bhi, Blo, bsc);
            NF_OK(hipGetLastError());
        }
        double dxq = 0.0, dxp = 0.0;
        const bool uni = uniform_spacing(pl.xq, &dxq);
        const bool uni_p = uniform_spacing(pl.xp, &dxp);
        ReduceArgs ra{};
        ra.G = G;
        ra.ldg = pl.N;
        ra.order = A.order.as<int>();
        ra.groups = A.groups.as<int2>();
        ra.uidx = A.uidx.as<int>();
        ra.usgn = A.usgn.as<float>();
        ra.nq = pl.nq;
        ra.kapq = d_kap + (size_t)pl.q * n_ang;
        ra.kap0 = d_kap + (size_t)f.axis * n_ang;
        ra.xq = xq;
        ra.wq = wq;
        ra.uniform = uni ? 1 : 0;
        ra.dxq = dxq;
        ra.x0 = f.c0 - center[f.axis];
        ra.k = k;
        ra.slot[0] = a1; ra.slot[1] = a2; ra.slot[2] = 3 + a1; ra.slot[3] = 3 + a2;
        ra.NL = lb[ln].nl;
        for (int u0 = 0; u0 < A.U; u0 += uchunk) {
            const int u1 = std::min(A.U, u0 + uchunk);
            const int mp = (int)round_up(2 * (size_t)(u1 - u0), GBM);
            const size_t total = (size_t)mp * pl.K;
            const int pb = (int)std::min<size_t>((total + 255) / 256, 16384);
            const bool fp16 = use_split16(mp, pl.N, split16);
            _Float16* Ahi = reinterpret_cast<_Float16*>(Am);
            _Float16* Alo = Ahi + total;
            if (fp16)
                hipLaunchKernelGGL(k_phase16, dim3(std::max(1, pb / 8)), dim3(256), 0, fs, Ahi, Alo, pl.K, total, u0, u1,
                                   pl.np, A.uval.as<double>(), xp, k, uni_p ? 1 : 0, dxp);
            else
                hipLaunchKernelGGL(k_phase, dim3(pb), dim3(256), 0, fs, Am, pl.K, total, u0, u1, pl.np,
                                   A.uval.as<double>(), xp, k);
            NF_OK(hipGetLastError());
            EventPair* ev = nullptr;
            NF_OK(next_events(&ev));
            gemm_ev.push_back(ev);
            NF_OK(hipEventRecord(ev->a, fs));
            if (fp16) {
                hipLaunchKernelGGL(k_hgemm3, dim3(pl.N / 128, mp / 128), dim3(256), 0, fs, Ahi, Alo, Bhi, Blo, G, mp, pl.N,
                                   pl.K, bsc);
                NF_OK(hipGetLastError());
            } else {
                NF_OK(launch_gemm(Am, Bm, G, mp, pl.N, pl.K, fs));
            }
            NF_OK(hipEventRecord(gemm_ev.back()->b, fs));
            gemm_split.push_back(fp16);
            const double fl = 2.0 * (2.0 * (u1 - u0)) * (8.0 * pl.nq) * pl.np;  // useful (unpadded) work
            flop += fl;
            if (fp16) flop_split += fl;
            ra.g0 = A.gstart[u0];
            ra.g1 = A.gstart[u1];
            ra.u0 = u0;
            const auto& M = C.mir[f.axis];
            const int nm = M.mstart[u1] - M.mstart[u0];
            if (mirror && 2 * (size_t)nm <= (size_t)(A.start[u1] - A.start[u0])) {
                // two or more angles per mirror group on average
                MirrorArgs ma{M.groups.as<int4>(), M.gkq.as<double>(), M.members.as<int>(), M.mstart[u0],
                              M.mstart[u1]};
                const dim3 grid((nm + MIRROR_GROUPS - 1) / MIRROR_GROUPS);
                if (ra.uniform)
                    hipLaunchKernelGGL(k_nf2ff_reduce_mirror<true>, grid, dim3(MIRROR_THREADS), 0, fs, ra, ma);
                else
                    hipLaunchKernelGGL(k_nf2ff_reduce_mirror<false>, grid, dim3(MIRROR_THREADS), 0, fs, ra, ma);
            } else {
                const int ng = ra.g1 - ra.g0;
                hipLaunchKernelGGL(k_nf2ff_reduce, dim3((ng + REDUCE_GROUPS - 1) / REDUCE_GROUPS),
                                   dim3(REDUCE_THREADS), 0, fs, ra);
            }
            NF_OK(hipGetLastError());
        }
    }
    if (L == 2) {
        NF_OK(hipEventRecord(C.ev_join, C.side));
        NF_OK(hipStreamWaitEvent(s, C.ev_join, 0));
    }
    typedef std::complex<double> cd;
    const cd fac = cd(0.0, k) * std::exp(cd(0.0, -k * r)) / (4.0 * M_PI * r);
    hipLaunchKernelGGL(k_nf2ff_project, dim3((unsigned)pblocks), dim3(256), 0, s, d_nl, L == 2 ? lb[1].nl : nullptr,
                       d_kap + 3 * n_ang,
                       d_kap + 3 * n_ang + n_theta, n_phi, n_ang, fac.real(), fac.imag(), r, d_ff, d_umax);
    NF_OK(hipGetLastError());
    NF_OK(hipEventRecord(all->b, s));
    const size_t hb = (4 * n_ang + pblocks + blocks) * sizeof(double);
    if (C.host.bytes < hb) NF_OK(C.host.alloc(hb));
    double* h = reinterpret_cast<double*>(C.host.p);
    NF_OK(hipMemcpyAsync(h, d_ff, (4 * n_ang + pblocks) * sizeof(double), hipMemcpyDeviceToHost, s));
    NF_OK(hipMemcpyAsync(h + 4 * n_ang + pblocks, d_part, blocks * sizeof(double), hipMemcpyDeviceToHost, s));
    NF_OK(hipStreamSynchronize(s));
    float ms = 0.f;
    NF_OK(hipEventElapsedTime(&ms, all->a, all->b));
    out->total_ms = ms;
    out->gemm_ms = 0.0;
    out->gemm_split_ms = 0.0;
    for (size_t q = 0; q < gemm_ev.size(); ++q) {
        NF_OK(hipEventElapsedTime(&ms, gemm_ev[q]->a, gemm_ev[q]->b));
        out->gemm_ms += ms;
        if (gemm_split[q]) out->gemm_split_ms += ms;
    }
    out->gemm_flop = flop;
    out->gemm_split_flop = flop_split;
    double prad = 0.0;
    for (size_t i = 0; i < blocks; ++i) prad += h[4 * n_ang + pblocks + i];
    out->prad = prad;
    double umax = 0.0;
    for (size_t i = 0; i < pblocks; ++i) umax = std::max(umax, h[4 * n_ang + i]);
    if (out->et) std::memcpy(out->et, h, 2 * n_ang * sizeof(double));
    if (out->ep) std::memcpy(out->ep, h + 2 * n_ang, 2 * n_ang * sizeof(double));
    double* dst[4] = {out->et_re, out->et_im, out->ep_re, out->ep_im};
    for (int q = 0; q < 4; ++q)
        if (dst[q]) {
            const double* src = h + (q / 2) * 2 * n_ang + (q % 2);
            for (size_t a = 0; a < n_ang; ++a) dst[q][a] = src[2 * a];
        }
    out->dmax = prad > 0.0 ? 4.0 * M_PI * umax / prad : 0.0;
    return hipSuccess;
}

}  // namespace oe
