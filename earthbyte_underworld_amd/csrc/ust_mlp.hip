// Learned Schur-preconditioner scaling: features, MLP inference, training (backward + Adam) and
// probing helpers, hand-written for gfx950.  The network maps the log-viscosity neighbourhood of
// a pressure cell (3^dim cells) to log10 of a factor on the "uw" diagonal approxS
// (UzawaPreconditionerTerm.c:195-212, used as the Q22 preconditioner, preconditioner.c:138-166).
// Weights are BF16, activations and accumulation FP32.  Hidden width is 64.
#include <math.h>

#include "ust_common.h"

#define MLP_H 64


// element viscosity from the K weights: sum_q WK[q] = sum_p w_p detJ eta_p (sum_q L_q = 1)
__device__ __forceinline__ double elem_logeta(const GridDev &g, const double *WK, int e) {
    const int nq = g.dim == 3 ? 27 : 9;
    double s = 0.0;
    for (int q = 0; q < nq; q++) s += WK[widx(g, e, q)];
    const double vol = g.dim == 3 ? g.hx * g.hy * g.hz : g.hx * g.hy;
    return log10(fmax(s / vol, 1e-300));
}

// feat[e*nin + k] = log10 eta(neighbour k) - log10 eta(e).  A neighbour outside the box reads as
// MLP_WALL: the walls act on a cell like a much stiffer neighbour, and the constant tells the network
// which faces are walls.
#define MLP_WALL 3.0f
__global__ void k_features(GridDev g, const double *__restrict__ WK, float *__restrict__ feat) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = e / (g.nx * g.ny);
    const int nin = g.dim == 3 ? 27 : 9;
    const double c = elem_logeta(g, WK, e);
    for (int t = 0; t < nin; t++) {
        const int di = t % 3 - 1, dj = (t / 3) % 3 - 1, dk = g.dim == 3 ? t / 9 - 1 : 0;
        const int ii = i + di, jj = j + dj, kk = k + dk;
        float v = MLP_WALL;
        if (ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz)
            v = (float)(elem_logeta(g, WK, ii + g.nx * (jj + g.ny * kk)) - c);
        feat[(size_t)e * nin + t] = v;
    }
}

// parameter layout: W1 [H][nin], W2 [H][H], W3 [H]; biases b1 [H], b2 [H], b3 [1]
template <int NIN>
struct MlpShared {
    float w1[MLP_H * NIN];
    float w2[MLP_H * MLP_H];
    float w3[MLP_H];
    float b1[MLP_H], b2[MLP_H], b3;
};

template <int NIN>
__device__ void load_params(MlpShared<NIN> &sm, const unsigned short *wbf, const float *bias) {
    for (int t = threadIdx.x; t < MLP_H * NIN; t += blockDim.x) sm.w1[t] = bf16_to_f32(wbf[t]);
    for (int t = threadIdx.x; t < MLP_H * MLP_H; t += blockDim.x) sm.w2[t] = bf16_to_f32(wbf[MLP_H * NIN + t]);
    for (int t = threadIdx.x; t < MLP_H; t += blockDim.x) {
        sm.w3[t] = bf16_to_f32(wbf[MLP_H * NIN + MLP_H * MLP_H + t]);
        sm.b1[t] = bias[t];
        sm.b2[t] = bias[MLP_H + t];
    }
    if (threadIdx.x == 0) sm.b3 = bias[2 * MLP_H];
    __syncthreads();
}

// forward for one sample; h1, h2 post-ReLU
template <int NIN>
__device__ __forceinline__ float mlp_fwd(const MlpShared<NIN> &sm, const float *x, float *h1, float *h2) {
#pragma unroll 4
    for (int j = 0; j < MLP_H; j++) {
        float s = sm.b1[j];
#pragma unroll
        for (int i = 0; i < NIN; i++) s += sm.w1[j * NIN + i] * x[i];
        h1[j] = fmaxf(s, 0.0f);
    }
#pragma unroll 4
    for (int j = 0; j < MLP_H; j++) {
        float s = sm.b2[j];
#pragma unroll 8
        for (int k = 0; k < MLP_H; k++) s += sm.w2[j * MLP_H + k] * h1[k];
        h2[j] = fmaxf(s, 0.0f);
    }
    float o = sm.b3;
#pragma unroll 8
    for (int j = 0; j < MLP_H; j++) o += sm.w3[j] * h2[j];
    return o;
}

template <int NIN>
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ feat, int n,
                                                     const unsigned short *__restrict__ wbf,
                                                     const float *__restrict__ bias, float *__restrict__ out) {
    __shared__ MlpShared<NIN> sm;
    load_params<NIN>(sm, wbf, bias);
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        float x[NIN], h1[MLP_H], h2[MLP_H];
#pragma unroll
        for (int i = 0; i < NIN; i++) x[i] = feat[(size_t)s * NIN + i];
        out[s] = mlp_fwd<NIN>(sm, x, h1, h2);
    }
}

// Gradients of L = 0.5/n sum (out - target)^2.  A block takes chunks of 64 samples: one thread per
// sample runs forward/backward into LDS, then the block turns the chunk into weight-gradient
// partials (each thread owns a fixed set of gradient entries, so no atomics); block partials are
// written to partial[block][nparam] and summed by k_mlp_reduce.
#define MLP_CHUNK 32
template <int NIN>
struct MlpTrainShared {
    MlpShared<NIN> p;
    float x[MLP_CHUNK][NIN];
    float h1[MLP_CHUNK][MLP_H];
    float h2[MLP_CHUNK][MLP_H];
    float dh1[MLP_CHUNK][MLP_H];
    float dh2[MLP_CHUNK][MLP_H];
    float dout[MLP_CHUNK];
    float lossacc[MLP_CHUNK];
};

template <int NIN>
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ feat, const float *__restrict__ target,
                                                  int n, const unsigned short *__restrict__ wbf,
                                                  const float *__restrict__ bias, float *__restrict__ partial,
                                                  float *__restrict__ loss_partial) {
    extern __shared__ unsigned char smem_raw[];
    MlpTrainShared<NIN> &sm = *reinterpret_cast<MlpTrainShared<NIN> *>(smem_raw);
    constexpr int NW = MLP_H * NIN + MLP_H * MLP_H + MLP_H;
    constexpr int NP = NW + 2 * MLP_H + 1;
    load_params<NIN>(sm.p, wbf, bias);
    // per-thread gradient accumulators for the entries t, t+256, ...
    constexpr int PER = (NP + 255) / 256;
    float gacc[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) gacc[r] = 0.0f;
    float lacc = 0.0f;
    const float inv_n = 1.0f / (float)n;
    const int nchunks = (n + MLP_CHUNK - 1) / MLP_CHUNK;
    for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int s0 = ch * MLP_CHUNK;
        const int ls = threadIdx.x;
        if (ls < MLP_CHUNK) {
            const int s = s0 + ls;
            float x[NIN], h1[MLP_H], h2[MLP_H];
            if (s < n) {
#pragma unroll
                for (int i = 0; i < NIN; i++) x[i] = feat[(size_t)s * NIN + i];
                const float o = mlp_fwd<NIN>(sm.p, x, h1, h2);
                const float d = (o - target[s]) * inv_n;
                lacc += 0.5f * (o - target[s]) * (o - target[s]) * inv_n;
                sm.dout[ls] = d;
                float dh2[MLP_H];
#pragma unroll 8
                for (int j = 0; j < MLP_H; j++) dh2[j] = h2[j] > 0.0f ? d * sm.p.w3[j] : 0.0f;
#pragma unroll 4
                for (int k = 0; k < MLP_H; k++) {
                    float a = 0.0f;
#pragma unroll 8
                    for (int j = 0; j < MLP_H; j++) a += sm.p.w2[j * MLP_H + k] * dh2[j];
                    sm.dh1[ls][k] = h1[k] > 0.0f ? a : 0.0f;
                }
#pragma unroll
                for (int i = 0; i < NIN; i++) sm.x[ls][i] = x[i];
#pragma unroll 8
                for (int j = 0; j < MLP_H; j++) {
                    sm.h1[ls][j] = h1[j];
                    sm.h2[ls][j] = h2[j];
                    sm.dh2[ls][j] = dh2[j];
                }
            } else {
                sm.dout[ls] = 0.0f;
                for (int i = 0; i < NIN; i++) sm.x[ls][i] = 0.0f;
                for (int j = 0; j < MLP_H; j++) {
                    sm.h1[ls][j] = 0.0f;
                    sm.h2[ls][j] = 0.0f;
                    sm.dh1[ls][j] = 0.0f;
                    sm.dh2[ls][j] = 0.0f;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int t = threadIdx.x + 256 * r;
            if (t >= NP) break;
            float a = 0.0f;
            if (t < MLP_H * NIN) {  // W1[j][i] = sum_s dh1[s][j] x[s][i]
                const int j = t / NIN, i = t % NIN;
                for (int s = 0; s < MLP_CHUNK; s++) a += sm.dh1[s][j] * sm.x[s][i];
            } else if (t < MLP_H * NIN + MLP_H * MLP_H) {  // W2[j][k] = sum_s dh2[s][j] h1[s][k]
                const int u = t - MLP_H * NIN, j = u / MLP_H, k = u % MLP_H;
                for (int s = 0; s < MLP_CHUNK; s++) a += sm.dh2[s][j] * sm.h1[s][k];
            } else if (t < NW) {  // W3[j] = sum_s dout[s] h2[s][j]
                const int j = t - MLP_H * NIN - MLP_H * MLP_H;
                for (int s = 0; s < MLP_CHUNK; s++) a += sm.dout[s] * sm.h2[s][j];
            } else if (t < NW + MLP_H) {
                const int j = t - NW;
                for (int s = 0; s < MLP_CHUNK; s++) a += sm.dh1[s][j];
            } else if (t < NW + 2 * MLP_H) {
                const int j = t - NW - MLP_H;
                for (int s = 0; s < MLP_CHUNK; s++) a += sm.dh2[s][j];
            } else {
                for (int s = 0; s < MLP_CHUNK; s++) a += sm.dout[s];
            }
            gacc[r] += a;
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int t = threadIdx.x + 256 * r;
        if (t < NP) partial[(size_t)blockIdx.x * NP + t] = gacc[r];
    }
    // block loss
    if (threadIdx.x < MLP_CHUNK) sm.lossacc[threadIdx.x] = lacc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float l = 0.0f;
        for (int s = 0; s < MLP_CHUNK; s++) l += sm.lossacc[s];
        loss_partial[blockIdx.x] = l;
    }
}

__global__ void k_mlp_reduce(const float *__restrict__ partial, int nblocks, int np, float *__restrict__ grad_w,
                             int nw, float *__restrict__ grad_b, const float *__restrict__ loss_partial,
                             float *__restrict__ loss) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < np) {
        float s = 0.0f;
        for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * np + t];
        if (t < nw) grad_w[t] = s;
        else grad_b[t - nw] = s;
    }
    if (t == 0 && loss) {
        float l = 0.0f;
        for (int b = 0; b < nblocks; b++) l += loss_partial[b];
        loss[0] = l;
    }
}

__global__ void k_mlp_adam(float *__restrict__ mw, unsigned short *__restrict__ wbf, float *__restrict__ bias,
                           const float *__restrict__ gw, const float *__restrict__ gb, float *__restrict__ m_w,
                           float *__restrict__ v_w, float *__restrict__ m_b, float *__restrict__ v_b, int nw, int nb,
                           float lr, float b1, float b2, float eps, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nw) {
        const float g = gw[t];
        const float m = b1 * m_w[t] + (1.0f - b1) * g, v = b2 * v_w[t] + (1.0f - b2) * g * g;
        m_w[t] = m;
        v_w[t] = v;
        const float w = mw[t] - lr * (m / c1) / (sqrtf(v / c2) + eps);
        mw[t] = w;
        wbf[t] = f32_to_bf16(w);
    } else if (t < nw + nb) {
        const int u = t - nw;
        const float g = gb[u];
        const float m = b1 * m_b[u] + (1.0f - b1) * g, v = b2 * v_b[u] + (1.0f - b2) * g * g;
        m_b[u] = m;
        v_b[u] = v;
        bias[u] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
}

__global__ void k_f32_to_bf16(const float *__restrict__ x, unsigned short *__restrict__ y, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = f32_to_bf16(x[t]);
}

// pc[e] = approxS[e] * 10^clip(out[e])
__global__ void k_apply_scaling(const double *__restrict__ approxS, const float *__restrict__ out,
                                double *__restrict__ pc, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double s = fmin(fmax((double)out[e], -4.0), 4.0);
    pc[e] = approxS[e] * pow(10.0, s);
}

// probing: v = 1 on the cells of colour c (3^dim colouring by (i%3, j%3, k%3)), 0 elsewhere
__global__ void k_color_indicator(GridDev g, int color, double *__restrict__ v) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = e / (g.nx * g.ny);
    const int c = (i % 3) + 3 * (j % 3) + (g.dim == 3 ? 9 * (k % 3) : 0);
    v[e] = (c == color) ? 1.0 : 0.0;
}

__global__ void k_color_gather(GridDev g, int color, const double *__restrict__ q, double *__restrict__ d) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int i = e % g.nx, j = (e / g.nx) % g.ny, k = e / (g.nx * g.ny);
    const int c = (i % 3) + 3 * (j % 3) + (g.dim == 3 ? 9 * (k % 3) : 0);
    if (c == color) d[e] = q[e];
}

__global__ void k_log_ratio(const double *__restrict__ d, const double *__restrict__ approxS,
                            float *__restrict__ t, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    t[e] = (float)log10(fmax(d[e], 1e-300) / approxS[e]);
}

extern "C" {

int ust_k_mlp_features(const ust_grid *gr, const double *WK, float *feat, void *stream) {
    GridDev g = grid_dev(gr);
    if (!WK || !feat) return UST_EINVAL;
    hipLaunchKernelGGL(k_features, dim3(ust_grid_size(g.nel, 128)), dim3(128), 0, ust_stream(stream), g, WK, feat);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_mlp_forward(const float *feat, int n, int nin, int hid, const unsigned short *wbf16, const float *bias,
                      float *out, float *act1, float *act2, void *stream) {
    (void)act1;
    (void)act2;
    if (hid != MLP_H || (nin != 27 && nin != 9) || n < 0 || !feat || !wbf16 || !bias || !out) return UST_EINVAL;
    if (n == 0) return UST_OK;
    unsigned grid = ust_grid_size(n, 256);
    if (grid > 2048) grid = 2048;
    if (nin == 27)
        hipLaunchKernelGGL(k_mlp_forward<27>, dim3(grid), dim3(256), 0, ust_stream(stream), feat, n, wbf16, bias, out);
    else
        hipLaunchKernelGGL(k_mlp_forward<9>, dim3(grid), dim3(256), 0, ust_stream(stream), feat, n, wbf16, bias, out);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

// grad_w must have room for nblocks*nparam partials after the nw gradient entries: the caller
// passes a workspace in act1 (>= 256 * nparam floats) and act2 (>= 256 floats) for the partials.
int ust_k_mlp_backward(const float *feat, const float *target, int n, int nin, int hid, const unsigned short *wbf16,
                       const float *bias, const float *act1, const float *act2, const float *out, float *grad_w,
                       float *grad_b, float *loss, void *stream) {
    (void)out;
    if (hid != MLP_H || (nin != 27 && nin != 9) || n <= 0 || !act1 || !act2) return UST_EINVAL;
    const int nw = MLP_H * nin + MLP_H * MLP_H + MLP_H, np = nw + 2 * MLP_H + 1;
    int nchunks = (n + MLP_CHUNK - 1) / MLP_CHUNK;
    int nblocks = nchunks < 256 ? nchunks : 256;
    float *partial = (float *)act1, *lpart = (float *)act2;
    hipStream_t st = ust_stream(stream);
    if (nin == 27)
        hipLaunchKernelGGL(k_mlp_grad<27>, dim3(nblocks), dim3(256), sizeof(MlpTrainShared<27>), st, feat, target, n,
                           wbf16, bias, partial, lpart);
    else
        hipLaunchKernelGGL(k_mlp_grad<9>, dim3(nblocks), dim3(256), sizeof(MlpTrainShared<9>), st, feat, target, n,
                           wbf16, bias, partial, lpart);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(ust_grid_size(np, 256)), dim3(256), 0, st, partial, nblocks, np, grad_w, nw,
                       grad_b, lpart, loss);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_mlp_adam(float *master_w, unsigned short *wbf16, float *bias, const float *grad_w, const float *grad_b,
                   float *m_w, float *v_w, float *m_b, float *v_b, int nw, int nb, float lr, float beta1, float beta2,
                   float eps, int step, void *stream) {
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(ust_grid_size(nw + nb, 256)), dim3(256), 0, ust_stream(stream), master_w,
                       wbf16, bias, grad_w, grad_b, m_w, v_w, m_b, v_b, nw, nb, lr, beta1, beta2, eps, c1, c2);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_mlp_to_bf16(const float *x, unsigned short *y, int n, void *stream) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(ust_grid_size(n, 256)), dim3(256), 0, ust_stream(stream), x, y, n);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_apply_scaling(const double *approxS, const float *out, double *pc, int n, void *stream) {
    hipLaunchKernelGGL(k_apply_scaling, dim3(ust_grid_size(n, 256)), dim3(256), 0, ust_stream(stream), approxS, out,
                       pc, n);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_color_indicator(const ust_grid *gr, int color, double *v, void *stream) {
    GridDev g = grid_dev(gr);
    hipLaunchKernelGGL(k_color_indicator, dim3(ust_grid_size(g.nel, 256)), dim3(256), 0, ust_stream(stream), g,
                       color, v);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_color_gather(const ust_grid *gr, int color, const double *q, double *d, void *stream) {
    GridDev g = grid_dev(gr);
    hipLaunchKernelGGL(k_color_gather, dim3(ust_grid_size(g.nel, 256)), dim3(256), 0, ust_stream(stream), g, color,
                       q, d);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_log_ratio(const double *d, const double *approxS, float *t, int n, void *stream) {
    hipLaunchKernelGGL(k_log_ratio, dim3(ust_grid_size(n, 256)), dim3(256), 0, ust_stream(stream), d, approxS, t, n);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

}  // extern "C"
