// oe_learn.hip -- the learned ringdown extrapolator: training and inference
// kernels for gfx950 and the C ABI behind oe_extrap_*.
//
// The network (1-D convolutions + MLP, about 2 M parameters) maps the early
// window of a port's V and I to 32 damped complex exponential modes whose sum
// continues the signal.  No counterpart in the reference; the decay test it
// shortens is src/fdtd/basic_engine.rs:128-141.
//
// Every matrix product -- forward and backward, convolutions as im2col
// products -- runs on the BF16 matrix cores (v_mfma_f32_32x32x8bf16_1k)
// with FP32 accumulation: both operands are bfloat16 (rounded to nearest
// even from float32), the results float32.  Master weights, biases, GELU, the
// mode synthesis and Adam are float32 (phases of the synthesis float64).
// Activations are kept channels-last, [batch, position, channel].

#include "oe_objects.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#pragma clang fp contract(off)

using namespace oe;

namespace {

constexpr int KW = 9, STRIDE = 2, PADW = 4;  // convolution kernel, stride, zero padding
constexpr float A_SCALE = 0.05f, F_MAX = 0.5f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------------ GEMM
// C(m, n) = sum_k A(m, k) B(k, n) [+ bias[n]] [+ C(m, n) when accumulate],
// A(m, k) = A[m sam + k sak], B(k, n) = B[k sbk + n sbn] in bfloat16,
// C(m, n) = C[m scm + n scn] float32.  64 x 64 tiles, 4 waves of 32 x 32,
// K steps of 32 staged in LDS as [row][k] so each lane reads its 4
// consecutive k of one row as one 8-byte load.  MFMA 32x32x8 bf16 operands:
// lane l supplies A[l % 32][4 (l / 32) .. +3] and B[4 (l / 32) .. +3][l % 32];
// accumulator register v of lane l holds C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
struct GemmArgs {
    const uint16_t* A;
    long sam, sak;
    const uint16_t* B;
    long sbk, sbn;
    float* C;
    long scm, scn;
    int M, N, K;
    const float* bias;
    int accumulate;
};

constexpr int TM = 64, TK = 32, TKP = TK + 4;

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    __shared__ __align__(16) uint16_t As[TM][TKP];
    __shared__ __align__(16) uint16_t Bs[TM][TKP];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int bm = blockIdx.y * TM, bn = blockIdx.x * TM;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    for (int k0 = 0; k0 < g.K; k0 += TK) {
#pragma unroll
        for (int e = 0; e < (TM * TK) / 256; ++e) {
            const int idx = tid + 256 * e;
            const int r = idx / TK, k = idx % TK;
            const int gm = bm + r, gn = bn + r, gk = k0 + k;
            As[r][k] = (gm < g.M && gk < g.K) ? g.A[gm * g.sam + gk * g.sak] : (uint16_t)0;
            Bs[r][k] = (gn < g.N && gk < g.K) ? g.B[gk * g.sbk + gn * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 8) {
            const int kr = kk + 4 * (lane >> 5);
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(&As[wm + (lane & 31)][kr]);
            const bf16x4 b = *reinterpret_cast<const bf16x4*>(&Bs[wn + (lane & 31)][kr]);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = bm + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = bn + wn + (lane & 31);
        if (row < g.M && col < g.N) {
            float* c = g.C + row * g.scm + col * g.scn;
            float val = acc[v];
            if (g.bias) val += g.bias[col];
            if (g.accumulate) val += *c;
            *c = val;
        }
    }
}

// -------------------------------------------------------------- elementwise
// cols[(b Lo + t) (C K) + c K + k] = x[b, 2 t + k - 4, c] (zero outside)
__global__ void k_im2col(const float* __restrict__ x, int B, int L, int C, uint16_t* __restrict__ cols) {
    const int Lo = L / STRIDE, CK = C * KW;
    const size_t n = (size_t)B * Lo * CK;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(q % CK);
        const size_t row = q / CK;
        const int t = (int)(row % Lo), b = (int)(row / Lo);
        const int c = j / KW, k = j % KW;
        const int s = STRIDE * t + k - PADW;
        cols[q] = (s >= 0 && s < L) ? f2bf(x[((size_t)b * L + s) * C + c]) : (uint16_t)0;
    }
}

// dx[b, s, c] = sum over (t, k) with 2 t + k - 4 = s of dcols[(b Lo + t), c K + k]
// (gather form of the adjoint: no atomics, fixed order t ascending)
__global__ void k_col2im(const float* __restrict__ dcols, int B, int L, int C, float* __restrict__ dx) {
    const int Lo = L / STRIDE, CK = C * KW;
    const size_t n = (size_t)B * L * C;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const size_t r = q / C;
        const int s = (int)(r % L), b = (int)(r / L);
        float acc = 0.f;
        const int t_lo = max(0, (s + PADW - (KW - 1) + STRIDE - 1) / STRIDE);
        const int t_hi = min(Lo - 1, (s + PADW) / STRIDE);
        for (int t = t_lo; t <= t_hi; ++t) {
            const int k = s + PADW - STRIDE * t;
            acc += dcols[((size_t)b * Lo + t) * CK + c * KW + k];
        }
        dx[q] = acc;
    }
}

__device__ __forceinline__ float gelu_f(float x) {
    const float c = 0.7978845608028654f;
    return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_d(float x) {
    const float c = 0.7978845608028654f;
    const float t = tanhf(c * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x * x);
}

__global__ void k_gelu(const float* __restrict__ pre, float* __restrict__ h, uint16_t* __restrict__ hb, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const float v = gelu_f(pre[q]);
        h[q] = v;
        hb[q] = f2bf(v);
    }
}

__global__ void k_gelu_back(const float* __restrict__ dh, const float* __restrict__ pre, float* __restrict__ dpre,
                            uint16_t* __restrict__ dpreb, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const float v = dh[q] * gelu_d(pre[q]);
        dpre[q] = v;
        dpreb[q] = f2bf(v);
    }
}

__global__ void k_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        b[q] = f2bf(a[q]);
}

// Column sums of X[M, N] in two deterministic passes: work-group (cb, r)
// sums rows r, r + R, r + 2R, ... of 64 columns with 4 row lanes, combined in
// a fixed order; then one thread per column adds the R partials in order.
constexpr int CS_R = 64;

__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ X, int M, int N,
                                                     float* __restrict__ part) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), rl = threadIdx.x >> 6;
    float s = 0.f;
    if (c < N)
        for (int r = blockIdx.y * 4 + rl; r < M; r += CS_R * 4) s += X[(size_t)r * N + c];
    red[rl][threadIdx.x & 63] = s;
    __syncthreads();
    if (rl == 0 && c < N) part[(size_t)blockIdx.y * N + c] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) +
                                                               red[2][threadIdx.x]) + red[3][threadIdx.x];
}

__global__ void k_colsum_final(const float* __restrict__ part, int N, float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (int r = 0; r < CS_R; ++r) s += part[(size_t)r * N + c];
    out[c] = s;
}

// Mode synthesis, loss and its gradient for one sample per work-group.
// o [B, 6 M]: per mode (softplus^-1 damping, logit frequency, V re, V im,
// I re, I im).  target [B, 2, T].  The residual goes to LDS (T <= TMAX),
// the loss partial to loss[b] (float64), dL/do to dout [B, 6 M]; scale is
// 1 / (B 2 T) of the mean.
constexpr int TMAX = 4096;

__global__ __launch_bounds__(256) void k_modes_loss(const float* __restrict__ o, const float* __restrict__ target,
                                                    int M, int T, double scale, double* __restrict__ loss,
                                                    float* __restrict__ dout, float* __restrict__ tails) {
    __shared__ float res[2][TMAX];
    __shared__ double red[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const float* ob = o + (size_t)b * 6 * M;
    double l = 0.0;
    for (int n = tid; n < T; n += blockDim.x) {
        double v = 0.0, i = 0.0;
        for (int m = 0; m < M; ++m) {
            const float* p = ob + 6 * m;
            const double alpha = (double)(log1pf(expf(-fabsf(p[0]))) + fmaxf(p[0], 0.f)) * A_SCALE;
            const double f = (double)F_MAX / (1.0 + exp(-(double)p[1]));
            const double mag = exp(-alpha * n);
            double sn, cs;
            sincos(2.0 * M_PI * f * n, &sn, &cs);
            v += mag * ((double)p[2] * cs - (double)p[3] * sn);
            i += mag * ((double)p[4] * cs - (double)p[5] * sn);
        }
        const float* tg = target + (size_t)b * 2 * T;
        const float rv = (float)(v - tg[n]), ri = (float)(i - tg[T + n]);
        res[0][n] = rv;
        res[1][n] = ri;
        if (tails) {
            tails[(size_t)b * 2 * T + n] = (float)v;
            tails[(size_t)b * 2 * T + T + n] = (float)i;
        }
        l += (double)rv * rv + (double)ri * ri;
    }
    red[tid] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) loss[b] = red[0];
    if (!dout) return;
    // d loss / d o for mode m = tid (M <= 256)
    if (tid < M) {
        const float* p = ob + 6 * tid;
        const double sp = (double)(log1pf(expf(-fabsf(p[0]))) + fmaxf(p[0], 0.f));
        const double alpha = sp * A_SCALE;
        const double sig = 1.0 / (1.0 + exp(-(double)p[1]));
        const double f = F_MAX * sig;
        double dvr = 0, dvi = 0, dir_ = 0, dii = 0, sr = 0, si = 0;
        for (int n = 0; n < T; ++n) {
            const double gv = 2.0 * scale * res[0][n], gi = 2.0 * scale * res[1][n];
            const double mag = exp(-alpha * n);
            double sn, cs;
            sincos(2.0 * M_PI * f * n, &sn, &cs);
            const double zr = mag * cs, zi = mag * sn;
            dvr += gv * zr;
            dvi -= gv * zi;
            dir_ += gi * zr;
            dii -= gi * zi;
            // a n z^n for V and I, weighted
            const double avr = (double)p[2] * zr - (double)p[3] * zi, avi = (double)p[2] * zi + (double)p[3] * zr;
            const double air = (double)p[4] * zr - (double)p[5] * zi, aii = (double)p[4] * zi + (double)p[5] * zr;
            sr += n * (gv * avr + gi * air);
            si += n * (gv * avi + gi * aii);
        }
        float* d = dout + (size_t)b * 6 * M + 6 * tid;
        const double sigmoid0 = 1.0 / (1.0 + exp(-(double)p[0]));
        d[0] = (float)(-sr * A_SCALE * sigmoid0);
        d[1] = (float)(2.0 * M_PI * (-si) * F_MAX * sig * (1.0 - sig));
        d[2] = (float)dvr;
        d[3] = (float)dvi;
        d[4] = (float)dir_;
        d[5] = (float)dii;
    }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const float gq = g[q];
        const float mq = b1 * m[q] + (1.0f - b1) * gq;
        const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
        m[q] = mq;
        v[q] = vq;
        w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
    }
}

int blocks_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

}  // namespace

// ------------------------------------------------------------------ host
struct oe_extrap {
    int device = 0;
    hipStream_t stream = nullptr;
    oe_extrap_dims d{};
    int C[4], L[4];
    size_t off[10], cnt[10], n_params = 0;
    DevBuf w, wb, g, m, v;         // fp32 master, bf16 copy, grads, Adam moments
    DevBuf act;                    // activations and scratch for the largest batch
    uint64_t adam_t = 0;
    // activation layout (element offsets into act, float32 unless noted)
    size_t x0, pre[3], h[3], hb[3], cols[3], dcols, dact, dpre, dpreb, flatb_dummy, pre4, h4, h4b, o, ob, dout, doutb,
        dh4, dpre4, dpre4b, loss, colpart, tails, bytes;
    ~oe_extrap() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

hipError_t colsum(hipStream_t s, const float* X, int M, int N, float* part, float* out) {
    hipLaunchKernelGGL(k_colsum_part, dim3((N + 63) / 64, CS_R), dim3(256), 0, s, X, M, N, part);
    hipLaunchKernelGGL(k_colsum_final, dim3((N + 255) / 256), dim3(256), 0, s, part, N, out);
    return hipGetLastError();
}

hipError_t gemm(hipStream_t s, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, float* C,
                long scm, long scn, int M, int N, int K, const float* bias, int accumulate) {
    GemmArgs g{A, sam, sak, B, sbk, sbn, C, scm, scn, M, N, K, bias, accumulate};
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + TM - 1) / TM, (M + TM - 1) / TM), dim3(256), 0, s, g);
    return hipGetLastError();
}

// Offsets in the activation block (bytes, 256-aligned) for batch B.
void layout(oe_extrap* x, size_t B) {
    size_t p = 0;
    auto take = [&](size_t bytes) {
        const size_t q = p;
        p += (bytes + 255) & ~(size_t)255;
        return q;
    };
    const int* C = x->C;
    const int* L = x->L;
    const int H = (int)x->d.hidden, O = 6 * (int)x->d.modes, T = (int)x->d.tail;
    x->x0 = take(B * L[0] * C[0] * 4);
    size_t maxcols = 0, maxact = 0;
    for (int l = 0; l < 3; ++l) {
        x->pre[l] = take(B * L[l + 1] * C[l + 1] * 4);
        x->h[l] = take(B * L[l + 1] * C[l + 1] * 4);
        x->hb[l] = take(B * L[l + 1] * C[l + 1] * 2);
        x->cols[l] = take(B * L[l + 1] * C[l] * KW * 2);
        maxcols = std::max(maxcols, B * L[l + 1] * (size_t)C[l] * KW);
        maxact = std::max(maxact, B * L[l] * (size_t)C[l]);
        maxact = std::max(maxact, B * L[l + 1] * (size_t)C[l + 1]);
    }
    x->dcols = take(maxcols * 4);
    x->dact = take(maxact * 4);
    x->dpre = take(maxact * 4);
    x->dpreb = take(maxact * 2);
    x->pre4 = take(B * H * 4);
    x->h4 = take(B * H * 4);
    x->h4b = take(B * H * 2);
    x->o = take(B * O * 4);
    x->dout = take(B * O * 4);
    x->doutb = take(B * O * 2);
    x->dh4 = take(B * H * 4);
    x->dpre4 = take(B * H * 4);
    x->dpre4b = take(B * H * 2);
    x->loss = take(B * 8);
    x->colpart = take((size_t)CS_R * std::max<size_t>({(size_t)H, (size_t)O, (size_t)C[1], (size_t)C[2], (size_t)C[3]}) * 4);
    x->tails = take(B * 2 * T * 4);
    x->bytes = p;
}

template <class T> T* at(oe_extrap* x, size_t off) { return reinterpret_cast<T*>(x->act.as<char>() + off); }

// Forward for B samples whose inputs are at x0; optional loss/backward.
oe_status run(oe_extrap* x, size_t B, const float* target_dev, bool backward, bool want_tails) {
    hipStream_t s = x->stream;
    const int* C = x->C;
    const int* L = x->L;
    const int H = (int)x->d.hidden, M = (int)x->d.modes, O = 6 * M, T = (int)x->d.tail;
    const float* W = x->w.as<float>();
    const uint16_t* Wb = x->wb.as<uint16_t>();
    float* G = x->g.as<float>();
    const float* in = at<float>(x, x->x0);
    for (int l = 0; l < 3; ++l) {
        const size_t rows = B * L[l + 1];
        const int CK = C[l] * KW;
        hipLaunchKernelGGL(k_im2col, dim3(blocks_for(rows * CK)), dim3(256), 0, s, in, (int)B, L[l], C[l],
                           at<uint16_t>(x, x->cols[l]));
        HIP_OK(gemm(s, at<uint16_t>(x, x->cols[l]), CK, 1, Wb + x->off[2 * l], 1, CK, at<float>(x, x->pre[l]), C[l + 1],
                    1, (int)rows, C[l + 1], CK, W + x->off[2 * l + 1], 0));
        const size_t n = rows * C[l + 1];
        hipLaunchKernelGGL(k_gelu, dim3(blocks_for(n)), dim3(256), 0, s, at<float>(x, x->pre[l]), at<float>(x, x->h[l]),
                           at<uint16_t>(x, x->hb[l]), n);
        in = at<float>(x, x->h[l]);
    }
    const int F = L[3] * C[3];
    HIP_OK(gemm(s, at<uint16_t>(x, x->hb[2]), F, 1, Wb + x->off[6], 1, F, at<float>(x, x->pre4), H, 1, (int)B, H, F,
                W + x->off[7], 0));
    hipLaunchKernelGGL(k_gelu, dim3(blocks_for(B * H)), dim3(256), 0, s, at<float>(x, x->pre4), at<float>(x, x->h4),
                       at<uint16_t>(x, x->h4b), B * H);
    HIP_OK(gemm(s, at<uint16_t>(x, x->h4b), H, 1, Wb + x->off[8], 1, H, at<float>(x, x->o), O, 1, (int)B, O, H,
                W + x->off[9], 0));
    if (!target_dev && !want_tails) return OE_OK;
    const double scale = 1.0 / ((double)B * 2.0 * T);
    hipLaunchKernelGGL(k_modes_loss, dim3((unsigned)B), dim3(256), 0, s, at<float>(x, x->o),
                       target_dev ? target_dev : at<float>(x, x->tails), M, T, scale, at<double>(x, x->loss),
                       backward ? at<float>(x, x->dout) : nullptr, want_tails ? at<float>(x, x->tails) : nullptr);
    HIP_OK(hipGetLastError());
    if (!backward) return OE_OK;
    // fc2
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(B * O)), dim3(256), 0, s, at<float>(x, x->dout),
                       at<uint16_t>(x, x->doutb), B * O);
    HIP_OK(gemm(s, at<uint16_t>(x, x->doutb), 1, O, at<uint16_t>(x, x->h4b), H, 1, G + x->off[8], H, 1, O, H, (int)B,
                nullptr, 0));
    HIP_OK(colsum(s, at<float>(x, x->dout), (int)B, O, at<float>(x, x->colpart), G + x->off[9]));
    HIP_OK(gemm(s, at<uint16_t>(x, x->doutb), O, 1, Wb + x->off[8], H, 1, at<float>(x, x->dh4), H, 1, (int)B, H, O,
                nullptr, 0));
    hipLaunchKernelGGL(k_gelu_back, dim3(blocks_for(B * H)), dim3(256), 0, s, at<float>(x, x->dh4),
                       at<float>(x, x->pre4), at<float>(x, x->dpre4), at<uint16_t>(x, x->dpre4b), B * H);
    // fc1
    HIP_OK(gemm(s, at<uint16_t>(x, x->dpre4b), 1, H, at<uint16_t>(x, x->hb[2]), F, 1, G + x->off[6], F, 1, H, F, (int)B,
                nullptr, 0));
    HIP_OK(colsum(s, at<float>(x, x->dpre4), (int)B, H, at<float>(x, x->colpart), G + x->off[7]));
    HIP_OK(gemm(s, at<uint16_t>(x, x->dpre4b), H, 1, Wb + x->off[6], F, 1, at<float>(x, x->dact), F, 1, (int)B, F, H,
                nullptr, 0));
    // convolutions, last to first: dact holds dL/dh of layer l
    for (int l = 2; l >= 0; --l) {
        const size_t rows = B * L[l + 1];
        const int CK = C[l] * KW, Co = C[l + 1];
        const size_t n = rows * Co;
        hipLaunchKernelGGL(k_gelu_back, dim3(blocks_for(n)), dim3(256), 0, s, at<float>(x, x->dact),
                           at<float>(x, x->pre[l]), at<float>(x, x->dpre), at<uint16_t>(x, x->dpreb), n);
        HIP_OK(gemm(s, at<uint16_t>(x, x->dpreb), 1, Co, at<uint16_t>(x, x->cols[l]), CK, 1, G + x->off[2 * l], CK, 1,
                    Co, CK, (int)rows, nullptr, 0));
        HIP_OK(colsum(s, at<float>(x, x->dpre), (int)rows, Co, at<float>(x, x->colpart), G + x->off[2 * l + 1]));
        if (l > 0) {
            HIP_OK(gemm(s, at<uint16_t>(x, x->dpreb), Co, 1, Wb + x->off[2 * l], CK, 1, at<float>(x, x->dcols), CK, 1,
                        (int)rows, CK, Co, nullptr, 0));
            hipLaunchKernelGGL(k_col2im, dim3(blocks_for(B * L[l] * C[l])), dim3(256), 0, s, at<float>(x, x->dcols),
                               (int)B, L[l], C[l], at<float>(x, x->dact));
        }
    }
    HIP_OK(hipGetLastError());
    return OE_OK;
}

oe_status upload_inputs(oe_extrap* x, const float* in, size_t B, const float* target, float** target_dev) {
    if (B == 0 || B > x->d.batch) return fail(OE_ERR_CONFIG, "batch %zu outside 1..%u", B, x->d.batch);
    HIP_OK(hipMemcpyAsync(at<float>(x, x->x0), in, B * x->L[0] * 2 * sizeof(float), hipMemcpyHostToDevice, x->stream));
    if (target) {
        HIP_OK(hipMemcpyAsync(at<float>(x, x->tails), target, B * 2 * x->d.tail * sizeof(float), hipMemcpyHostToDevice,
                              x->stream));
        *target_dev = at<float>(x, x->tails);
    }
    return OE_OK;
}

double sum_loss(oe_extrap* x, size_t B) {
    std::vector<double> l(B);
    (void)hipMemcpyAsync(l.data(), at<double>(x, x->loss), B * 8, hipMemcpyDeviceToHost, x->stream);
    (void)hipStreamSynchronize(x->stream);
    double s = 0.0;
    for (double v : l) s += v;
    return s / ((double)B * 2.0 * x->d.tail);
}

}  // namespace

extern "C" {

size_t oe_extrap_param_count(const oe_extrap* x) { return x ? x->n_params : 0; }

oe_status oe_extrap_create(int device, const oe_extrap_dims* dims, const float* params, uint64_t seed,
                           oe_extrap** out) {
    if (!dims || !out) return fail(OE_ERR_CONFIG, "null argument");
    const oe_extrap_dims& d = *dims;
    if (d.L0 < 8 || d.L0 % 8 || !d.c1 || !d.c2 || !d.c3 || !d.hidden || !d.modes || d.modes > 256 || !d.batch ||
        !d.tail || d.tail > (uint32_t)TMAX)
        return fail(OE_ERR_CONFIG, "bad extrapolator dimensions (L0 multiple of 8, modes <= 256, tail <= %d)", TMAX);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(OE_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(OE_ERR_CONFIG, "device %d out of range", device);
    std::unique_ptr<oe_extrap> x(new oe_extrap());
    x->device = device;
    x->d = d;
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
    x->C[0] = 2; x->C[1] = (int)d.c1; x->C[2] = (int)d.c2; x->C[3] = (int)d.c3;
    x->L[0] = (int)d.L0;
    for (int l = 1; l < 4; ++l) x->L[l] = x->L[l - 1] / STRIDE;
    const size_t F = (size_t)x->L[3] * x->C[3];
    const size_t shapes[10] = {(size_t)x->C[1] * x->C[0] * KW, (size_t)x->C[1], (size_t)x->C[2] * x->C[1] * KW,
                               (size_t)x->C[2], (size_t)x->C[3] * x->C[2] * KW, (size_t)x->C[3], d.hidden * F,
                               d.hidden, 6 * (size_t)d.modes * d.hidden, 6 * (size_t)d.modes};
    size_t p = 0;
    for (int q = 0; q < 10; ++q) {
        x->off[q] = p;
        x->cnt[q] = shapes[q];
        p += shapes[q];
    }
    x->n_params = p;
    std::vector<float> w(p, 0.f);
    if (params) {
        std::memcpy(w.data(), params, p * sizeof(float));
    } else {  // He initialisation of the weights, zero biases
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        const size_t fan[5] = {(size_t)x->C[0] * KW, (size_t)x->C[1] * KW, (size_t)x->C[2] * KW, F, d.hidden};
        for (int l = 0; l < 5; ++l)
            for (size_t q = 0; q < x->cnt[2 * l]; ++q)
                w[x->off[2 * l] + q] = (float)(nd(rng) * std::sqrt(2.0 / (double)fan[l]));
    }
    HIP_OK(x->w.alloc(p * 4));
    HIP_OK(x->wb.alloc(p * 2));
    HIP_OK(x->g.alloc(p * 4));
    HIP_OK(x->m.alloc(p * 4));
    HIP_OK(x->v.alloc(p * 4));
    HIP_OK(hipMemcpy(x->w.p, w.data(), p * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(x->m.p, 0, p * 4));
    HIP_OK(hipMemset(x->v.p, 0, p * 4));
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(p)), dim3(256), 0, x->stream, x->w.as<float>(), x->wb.as<uint16_t>(), p);
    HIP_OK(hipGetLastError());
    layout(x.get(), d.batch);
    HIP_OK(x->act.alloc(x->bytes));
    HIP_OK(hipMemsetAsync(x->act.p, 0, x->bytes, x->stream));
    HIP_OK(hipStreamSynchronize(x->stream));
    *out = x.release();
    return OE_OK;
}

void oe_extrap_destroy(oe_extrap* x) { delete x; }

oe_status oe_extrap_params(oe_extrap* x, float* out, size_t n) {
    if (!x || !out) return fail(OE_ERR_CONFIG, "null argument");
    if (n < x->n_params) return fail(OE_ERR_CONFIG, "buffer holds %zu of %zu parameters", n, x->n_params);
    HIP_OK(hipSetDevice(x->device));
    HIP_OK(hipStreamSynchronize(x->stream));
    HIP_OK(hipMemcpy(out, x->w.p, x->n_params * 4, hipMemcpyDeviceToHost));
    return OE_OK;
}

oe_status oe_extrap_set_params(oe_extrap* x, const float* in, size_t n) {
    if (!x || !in) return fail(OE_ERR_CONFIG, "null argument");
    if (n != x->n_params) return fail(OE_ERR_CONFIG, "expected %zu parameters, got %zu", x->n_params, n);
    HIP_OK(hipSetDevice(x->device));
    HIP_OK(hipMemcpyAsync(x->w.p, in, n * 4, hipMemcpyHostToDevice, x->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, x->stream, x->w.as<float>(), x->wb.as<uint16_t>(), n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(x->stream));
    return OE_OK;
}

oe_status oe_extrap_forward(oe_extrap* x, const float* inputs, size_t B, float* modes, float* tails) {
    if (!x || !inputs) return fail(OE_ERR_CONFIG, "null argument");
    HIP_OK(hipSetDevice(x->device));
    OE_TRY(upload_inputs(x, inputs, B, nullptr, nullptr));
    OE_TRY(run(x, B, nullptr, false, tails != nullptr));
    const size_t O = 6 * (size_t)x->d.modes;
    if (modes) HIP_OK(hipMemcpyAsync(modes, at<float>(x, x->o), B * O * 4, hipMemcpyDeviceToHost, x->stream));
    if (tails)
        HIP_OK(hipMemcpyAsync(tails, at<float>(x, x->tails), B * 2 * x->d.tail * 4, hipMemcpyDeviceToHost, x->stream));
    HIP_OK(hipStreamSynchronize(x->stream));
    return OE_OK;
}

oe_status oe_extrap_gradients(oe_extrap* x, const float* inputs, const float* target, size_t B, double* loss,
                              float* grads) {
    if (!x || !inputs || !target) return fail(OE_ERR_CONFIG, "null argument");
    HIP_OK(hipSetDevice(x->device));
    float* tdev = nullptr;
    OE_TRY(upload_inputs(x, inputs, B, target, &tdev));
    OE_TRY(run(x, B, tdev, true, false));
    if (loss) *loss = sum_loss(x, B);
    if (grads) HIP_OK(hipMemcpyAsync(grads, x->g.p, x->n_params * 4, hipMemcpyDeviceToHost, x->stream));
    HIP_OK(hipStreamSynchronize(x->stream));
    return OE_OK;
}

oe_status oe_extrap_train_step(oe_extrap* x, const float* inputs, const float* target, size_t B, float lr,
                               double* loss) {
    if (!x || !inputs || !target) return fail(OE_ERR_CONFIG, "null argument");
    if (!(lr > 0.f)) return fail(OE_ERR_CONFIG, "learning rate must be positive");
    HIP_OK(hipSetDevice(x->device));
    float* tdev = nullptr;
    OE_TRY(upload_inputs(x, inputs, B, target, &tdev));
    OE_TRY(run(x, B, tdev, true, false));
    x->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)x->adam_t), c2 = 1.0f - std::pow(b2, (float)x->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(x->n_params)), dim3(256), 0, x->stream, x->w.as<float>(),
                       x->g.as<float>(), x->m.as<float>(), x->v.as<float>(), x->n_params, lr, b1, b2, 1e-8f, c1, c2);
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(x->n_params)), dim3(256), 0, x->stream, x->w.as<float>(),
                       x->wb.as<uint16_t>(), x->n_params);
    HIP_OK(hipGetLastError());
    if (loss) *loss = sum_loss(x, B);
    HIP_OK(hipStreamSynchronize(x->stream));
    return OE_OK;
}

}  // extern "C"
