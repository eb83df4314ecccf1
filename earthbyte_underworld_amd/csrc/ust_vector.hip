// Krylov / smoother vector kernels (the PETSc Vec operations inside the BSSCR CG and the
// multigrid smoothers).  Scalars that the device computes (dots, CG ratios) stay in device
// memory, so a Krylov iteration issues no host round-trip except its convergence check.
#include "ust_common.h"

#define GRID_STRIDE(i, n) \
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

static inline unsigned vgrid(size_t n) {
    size_t g = (n + UST_BLOCK - 1) / UST_BLOCK;
    if (g > 8192) g = 8192;
    return (unsigned)(g == 0 ? 1 : g);
}

template <typename T>
__global__ void kv_fill(T *x, T v, size_t n) {
    GRID_STRIDE(i, n) x[i] = v;
}
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    // splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__global__ void kv_hash(double *x, size_t n, unsigned long long seed) {
    GRID_STRIDE(i, n) {
        const unsigned long long h = mix64(seed * 0x100000001B3ULL + i);
        x[i] = 2.0 * ((double)(h >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}
__global__ void kv_sign(double *x, size_t n, unsigned long long seed) {
    GRID_STRIDE(i, n) x[i] = (mix64(seed * 0x100000001B3ULL + i) >> 63) ? 1.0 : -1.0;
}
__global__ void kv_d2f(const double *x, float *y, size_t n) {
    GRID_STRIDE(i, n) y[i] = (float)x[i];
}
__global__ void kv_d2bf(const double *x, unsigned short *y, size_t n) {
    GRID_STRIDE(i, n) y[i] = f32_to_bf16((float)x[i]);
}
__global__ void kv_sw_step(double *pfw, const double *inc, double dt, size_t n) {
    GRID_STRIDE(i, n) {
        const double v = pfw[i] + dt * inc[i];
        pfw[i] = v < 0.0 ? 0.0 : v;
    }
}
__global__ void kv_permute(const unsigned *perm, const double *src, double *dst, size_t n) {
    GRID_STRIDE(i, n) dst[i] = src[perm[i]];
}
__global__ void kv_f2d(const float *x, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] = (double)x[i];
}
__global__ void kv_axpy(double a, const double *x, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] += a * x[i];
}
__global__ void kv_axpby(double a, const double *x, double b, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] = a * x[i] + b * y[i];
}
__global__ void kv_addf(const float *x, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] += (double)x[i];
}
__global__ void kv_sub(const double *a, const double *b, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] = a[i] - b[i];
}
__global__ void kv_mul(const double *a, const double *b, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] = a[i] * b[i];
}
__global__ void kv_div(const double *a, const double *b, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] = a[i] / b[i];
}

// two-stage deterministic reductions: stage 1 writes one partial per block (<= 1024 blocks)
template <typename T, bool SQ>
__global__ __launch_bounds__(256) void kv_dot_stage1(const T *x, const T *y, size_t n, double *partial) {
    __shared__ double sh[256 / 64];
    double s = 0.0;
    GRID_STRIDE(i, n) s += SQ ? (double)x[i] * (double)y[i] : (double)x[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w];
        partial[blockIdx.x] = t;
    }
}

// x . y with x in FP64 and y in FP32 (the PCG's r . z with the FP32 V-cycle output z)
__global__ __launch_bounds__(256) void kv_dot_mixed_stage1(const double *x, const float *y, size_t n,
                                                           double *partial) {
    __shared__ double sh[256 / 64];
    double s = 0.0;
    GRID_STRIDE(i, n) s += x[i] * (double)y[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// sum of the stage-1 partials (one 256-thread block); the total is valid in thread 0
__device__ __forceinline__ double block_sum_partials(const double *partial, int np) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += partial[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void kv_reduce_stage2(const double *partial, int np, double *out) {
    const double s = block_sum_partials(partial, np);
    if (threadIdx.x == 0) out[0] = s;
}

// stage 2 of r . z fused with the CG beta update (kv_cg_beta)
__global__ __launch_bounds__(256) void kv_reduce_beta(const double *partial, int np, double *rz, double *rz_old,
                                                      double *c) {
    const double s = block_sum_partials(partial, np);
    if (threadIdx.x != 0) return;
    rz[0] = s;
    if (c[1] != 0.0) return;
    c[3] = s / rz_old[0];
    rz_old[0] = s;
}

// stage 2 of r . r fused with the CG convergence check (kv_cg_check)
__global__ __launch_bounds__(256) void kv_reduce_check(const double *partial, int np, double *rr, double *c) {
    const double s = block_sum_partials(partial, np);
    if (threadIdx.x != 0) return;
    rr[0] = s;
    if (c[1] != 0.0) return;
    c[2] += 1.0;
    if (!(s == s)) c[1] = 2.0;
    else if (s <= c[0]) c[1] = 1.0;
}

__global__ __launch_bounds__(256) void kv_absmax_stage1(const double *x, size_t n, double *partial) {
    __shared__ double sh[4];
    double s = 0.0;
    GRID_STRIDE(i, n) s = fmax(s, fabs(x[i]));
    for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ __launch_bounds__(256) void kv_max_stage2(const double *partial, int np, double *out) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s = fmax(s, partial[i]);
    for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

static inline unsigned rgrid(size_t n) {
    size_t g = (n + UST_BLOCK * 4 - 1) / (UST_BLOCK * 4);
    if (g > 1024) g = 1024;
    return (unsigned)(g == 0 ? 1 : g);
}

__global__ void kv_mask_zero(double *x, const unsigned char *mask, size_t n) {
    GRID_STRIDE(i, n) if (mask[i]) x[i] = 0.0;
}

__global__ void kv_remove_mean(double *x, size_t n, const double *sum) {
    const double m = sum[0] / (double)n;
    GRID_STRIDE(i, n) x[i] -= m;
}

template <typename T>
__global__ void kv_cg_p(const T *z, T *p, const double *rz_new, const double *rz_old, int first, size_t n) {
    const T beta = first ? (T)0 : (T)(rz_new[0] / rz_old[0]);
    GRID_STRIDE(i, n) p[i] = z[i] + (first ? (T)0 : beta * p[i]);
}

template <typename T>
__global__ void kv_cg_xr(T *x, T *r, const T *p, const T *q, const double *num, const double *den, size_t n) {
    const T a = (T)(num[0] / den[0]);
    GRID_STRIDE(i, n) {
        x[i] += a * p[i];
        r[i] -= a * q[i];
    }
}

// Device-side CG control block c: c[0] = stopping threshold on r.r, c[1] = state (0 running,
// 1 converged, 2 NaN), c[2] = iterations taken, c[3] = beta.  Once the state leaves 0 every update
// kernel below returns at once, so iterations launched past convergence change nothing.
__global__ void kv_cg_start(const double *bb, double rtol2, double *c, double *rz_old) {
    c[0] = rtol2 * bb[0];
    c[1] = bb[0] > 0.0 ? 0.0 : 1.0;
    c[2] = 0.0;
    c[3] = 0.0;
    rz_old[0] = 1.0;
}

__global__ void kv_cg_beta(const double *rz, double *rz_old, double *c) {
    if (c[1] != 0.0) return;
    c[3] = rz[0] / rz_old[0];
    rz_old[0] = rz[0];
}

__global__ void kv_cg_p_guarded(const double *z, double *p, const double *c, size_t n) {
    if (c[1] != 0.0) return;
    const double beta = c[3];
    GRID_STRIDE(i, n) p[i] = z[i] + beta * p[i];
}

__global__ void kv_cg_xr_guarded(double *x, double *r, const double *p, const double *q, const double *rz,
                                 const double *pq, const double *c, size_t n) {
    if (c[1] != 0.0) return;
    const double a = rz[0] / pq[0];
    GRID_STRIDE(i, n) {
        x[i] += a * p[i];
        r[i] -= a * q[i];
    }
}

__global__ void kv_cg_p_guarded_zf(const float *z, double *p, const double *c, size_t n) {
    if (c[1] != 0.0) return;
    const double beta = c[3];
    GRID_STRIDE(i, n) p[i] = (double)z[i] + beta * p[i];
}

// x += a p, r -= a q and the block partials of r . r (zero partials once converged, so the stage-2 sum is 0
// and the check kernel ignores it)
__global__ __launch_bounds__(256) void kv_cg_xr_rr_guarded(double *x, double *r, const double *p, const double *q,
                                                           const double *rz, const double *pq, const double *c,
                                                           size_t n, double *partial) {
    __shared__ double sh[256 / 64];
    double s = 0.0;
    if (c[1] == 0.0) {
        const double a = rz[0] / pq[0];
        GRID_STRIDE(i, n) {
            x[i] += a * p[i];
            const double ri = r[i] - a * q[i];
            r[i] = ri;
            s += ri * ri;
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// The FP32 velocity PCG inside FP64 iterative refinement: p = z + beta p, and x += a p, r -= a q with the
// block partials of r . r accumulated in FP64
__global__ void kv_cg_p_guarded_f32(const float *z, float *p, const double *c, size_t n) {
    if (c[1] != 0.0) return;
    const float beta = (float)c[3];
    GRID_STRIDE(i, n) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(256) void kv_cg_xr_rr_guarded_f32(float *x, float *r, const float *p, const float *q,
                                                               const double *rz, const double *pq, const double *c,
                                                               size_t n, double *partial) {
    __shared__ double sh[256 / 64];
    double s = 0.0;
    if (c[1] == 0.0) {
        const float a = (float)(rz[0] / pq[0]);
        GRID_STRIDE(i, n) {
            x[i] += a * p[i];
            const float ri = r[i] - a * q[i];
            r[i] = ri;
            s += (double)ri * (double)ri;
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// The FP32 PCG's streaming kernels on 16-byte aligned vectors: four dofs per thread and step (one 16-byte load
// or store per vector), restrict-qualified so the loads of a step are issued together; the n % 4 tail dofs are
// taken by the first threads after the grid-stride loop.  The sums run in a fixed order per thread.
__device__ __forceinline__ size_t gid() { return blockIdx.x * (size_t)blockDim.x + threadIdx.x; }
__device__ __forceinline__ size_t gstride() { return (size_t)gridDim.x * blockDim.x; }

__global__ __launch_bounds__(256) void kv4_cg_xr_rr_guarded_f32(float *__restrict__ x, float *__restrict__ r,
                                                                const float *__restrict__ p,
                                                                const float *__restrict__ q, const double *rz,
                                                                const double *pq, const double *c, size_t n,
                                                                double *partial) {
    __shared__ double sh[256 / 64];
    double s = 0.0;
    if (c[1] == 0.0) {
        const float a = (float)(rz[0] / pq[0]);
        const size_t n4 = n / 4;
        float4 *x4 = (float4 *)x, *r4 = (float4 *)r;
        const float4 *p4 = (const float4 *)p, *q4 = (const float4 *)q;
        for (size_t i = gid(); i < n4; i += gstride()) {
            float4 xv = x4[i], rv = r4[i];
            const float4 pv = p4[i], qv = q4[i];
            xv.x += a * pv.x;
            xv.y += a * pv.y;
            xv.z += a * pv.z;
            xv.w += a * pv.w;
            rv.x = rv.x - a * qv.x;
            rv.y = rv.y - a * qv.y;
            rv.z = rv.z - a * qv.z;
            rv.w = rv.w - a * qv.w;
            x4[i] = xv;
            r4[i] = rv;
            s += (double)rv.x * (double)rv.x;
            s += (double)rv.y * (double)rv.y;
            s += (double)rv.z * (double)rv.z;
            s += (double)rv.w * (double)rv.w;
        }
        for (size_t i = 4 * n4 + gid(); i < n; i += gstride()) {
            x[i] += a * p[i];
            const float ri = r[i] - a * q[i];
            r[i] = ri;
            s += (double)ri * (double)ri;
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void kv4_dot_f32(const float *__restrict__ x, const float *__restrict__ y, size_t n,
                                                   double *partial) {
    __shared__ double sh[256 / 64];
    double s = 0.0;
    const size_t n4 = n / 4;
    const float4 *x4 = (const float4 *)x, *y4 = (const float4 *)y;
    for (size_t i = gid(); i < n4; i += gstride()) {
        const float4 a = x4[i], b = y4[i];
        s += (double)a.x * (double)b.x;
        s += (double)a.y * (double)b.y;
        s += (double)a.z * (double)b.z;
        s += (double)a.w * (double)b.w;
    }
    for (size_t i = 4 * n4 + gid(); i < n; i += gstride()) s += (double)x[i] * (double)y[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void kv4_cg_p_guarded_f32(const float *__restrict__ z, float *__restrict__ p,
                                                            const double *c, size_t n) {
    if (c[1] != 0.0) return;
    const float beta = (float)c[3];
    const size_t n4 = n / 4;
    const float4 *z4 = (const float4 *)z;
    float4 *p4 = (float4 *)p;
    for (size_t i = gid(); i < n4; i += gstride()) {
        const float4 zv = z4[i];
        float4 pv = p4[i];
        pv.x = zv.x + beta * pv.x;
        pv.y = zv.y + beta * pv.y;
        pv.z = zv.z + beta * pv.z;
        pv.w = zv.w + beta * pv.w;
        p4[i] = pv;
    }
    for (size_t i = 4 * n4 + gid(); i < n; i += gstride()) p[i] = z[i] + beta * p[i];
}

static inline bool al16(const void *a) { return ((uintptr_t)a & 15u) == 0; }

// y = (float)(s x): the FP64 residual scaled to unit norm for the FP32 correction solve
__global__ void kv_d2f_scaled(const double *x, float *y, double s, size_t n) {
    GRID_STRIDE(i, n) y[i] = (float)(s * x[i]);
}

// y += s x with x in FP32: the correction of a refinement step
__global__ void kv_axpy_f32_to_f64(double s, const float *x, double *y, size_t n) {
    GRID_STRIDE(i, n) y[i] += s * (double)x[i];
}

__global__ void kv_cg_check(const double *rr, double *c) {
    if (c[1] != 0.0) return;
    c[2] += 1.0;
    if (!(rr[0] == rr[0])) c[1] = 2.0;
    else if (rr[0] <= c[0]) c[1] = 1.0;
}

template <typename T>
__global__ void kv_cheby(const T *b, const T *Ax, const T *invd, T *d, T *x, T c1, T c2, int first, size_t n) {
    GRID_STRIDE(i, n) {
        const T r = b[i] - Ax[i];
        const T dn = (first ? (T)0 : c1 * d[i]) + c2 * invd[i] * r;
        d[i] = dn;
        x[i] += dn;
    }
}

// first Chebyshev step from x = 0: d = c2 invd b, x = d
template <typename T>
__global__ void kv_cheby_first(const T *b, const T *invd, T *d, T *x, T c2, size_t n) {
    GRID_STRIDE(i, n) {
        const T dn = c2 * invd[i] * b[i];
        d[i] = dn;
        x[i] = dn;
    }
}

// first Chebyshev step of the FP32 V-cycle on an FP64 right-hand side: b32 = (float) b, d = c2 invd b32, x = d
__global__ void kv_cheby_first_from64(const double *b, float *b32, const float *invd, float *d, float *x, float c2,
                                      size_t n) {
    GRID_STRIDE(i, n) {
        const float bi = (float)b[i];
        const float dn = c2 * invd[i] * bi;
        b32[i] = bi;
        d[i] = dn;
        x[i] = dn;
    }
}

template <typename T>
__global__ void kv_residual(const T *b, const T *Ax, T *r, size_t n) {
    GRID_STRIDE(i, n) r[i] = b[i] - Ax[i];
}

#define LV(kern, n, ...)                                                                               \
    do {                                                                                               \
        hipLaunchKernelGGL(kern, dim3(vgrid(n)), dim3(UST_BLOCK), 0, ust_stream(stream), __VA_ARGS__); \
        UST_LAUNCH_CHECK();                                                                            \
        return UST_OK;                                                                                 \
    } while (0)

extern "C" {

int ust_k_fill_f64(double *x, double v, size_t n, void *stream) { LV(kv_fill<double>, n, x, v, n); }
int ust_k_fill_f32(float *x, float v, size_t n, void *stream) { LV(kv_fill<float>, n, x, v, n); }
int ust_k_fill_hash_f64(double *x, size_t n, unsigned long long seed, void *stream) {
    LV(kv_hash, n, x, n, seed);
}
int ust_k_fill_sign_f64(double *x, size_t n, unsigned long long seed, void *stream) {
    LV(kv_sign, n, x, n, seed);
}
int ust_k_copy_f64_to_f32(const double *x, float *y, size_t n, void *stream) { LV(kv_d2f, n, x, y, n); }
int ust_k_copy_f32_to_f64(const float *x, double *y, size_t n, void *stream) { LV(kv_f2d, n, x, y, n); }
int ust_k_copy_f64_to_bf16(const double *x, unsigned short *y, size_t n, void *stream) { LV(kv_d2bf, n, x, y, n); }
int ust_k_strain_weakening_step(double *pfw, const double *pfw_inc, int np, double dt, void *stream) {
    if (np < 0 || (np > 0 && (!pfw || !pfw_inc))) return UST_EINVAL;
    LV(kv_sw_step, (size_t)np, pfw, pfw_inc, dt, (size_t)np);
}
int ust_k_permute_f64(const unsigned *perm, const double *src, double *dst, int n, void *stream) {
    if (n < 0 || (n > 0 && (!perm || !src || !dst))) return UST_EINVAL;
    LV(kv_permute, (size_t)n, perm, src, dst, (size_t)n);
}
int ust_k_axpy_f64(double a, const double *x, double *y, size_t n, void *stream) { LV(kv_axpy, n, a, x, y, n); }
int ust_k_axpby_f64(double a, const double *x, double b, double *y, size_t n, void *stream) {
    LV(kv_axpby, n, a, x, b, y, n);
}
int ust_k_add_f32_to_f64(const float *x, double *y, size_t n, void *stream) { LV(kv_addf, n, x, y, n); }
int ust_k_sub_f64(const double *a, const double *b, double *y, size_t n, void *stream) { LV(kv_sub, n, a, b, y, n); }
int ust_k_scale_mul_f64(const double *a, const double *b, double *y, size_t n, void *stream) {
    LV(kv_mul, n, a, b, y, n);
}
int ust_k_div_f64(const double *a, const double *b, double *y, size_t n, void *stream) { LV(kv_div, n, a, b, y, n); }

int ust_k_dot_f64(const double *x, const double *y, size_t n, double *partial, double *out, void *stream) {
    const unsigned g = rgrid(n);
    hipLaunchKernelGGL((kv_dot_stage1<double, true>), dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    hipLaunchKernelGGL(kv_reduce_stage2, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, out);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_dot_f32(const float *x, const float *y, size_t n, double *partial, double *out, void *stream) {
    const unsigned g = rgrid(n);
    if (al16(x) && al16(y))
        hipLaunchKernelGGL(kv4_dot_f32, dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    else
        hipLaunchKernelGGL((kv_dot_stage1<float, true>), dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    hipLaunchKernelGGL(kv_reduce_stage2, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, out);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_sum_f64(const double *x, size_t n, double *partial, double *out, void *stream) {
    const unsigned g = rgrid(n);
    hipLaunchKernelGGL((kv_dot_stage1<double, false>), dim3(g), dim3(256), 0, ust_stream(stream), x, x, n, partial);
    hipLaunchKernelGGL(kv_reduce_stage2, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, out);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_absmax_f64(const double *x, size_t n, double *partial, double *out, void *stream) {
    const unsigned g = rgrid(n);
    hipLaunchKernelGGL(kv_absmax_stage1, dim3(g), dim3(256), 0, ust_stream(stream), x, n, partial);
    hipLaunchKernelGGL(kv_max_stage2, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, out);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_mask_zero_f64(double *x, const unsigned char *mask, size_t n, void *stream) {
    LV(kv_mask_zero, n, x, mask, n);
}

int ust_k_remove_mean_f64(double *x, size_t n, const double *sum, void *stream) {
    LV(kv_remove_mean, n, x, n, sum);
}

int ust_k_cg_update_p_f64(const double *z, double *p, const double *rz_new, const double *rz_old, int first, size_t n,
                          void *stream) {
    LV(kv_cg_p<double>, n, z, p, rz_new, rz_old, first, n);
}
int ust_k_cg_update_p_f32(const float *z, float *p, const double *rz_new, const double *rz_old, int first, size_t n,
                          void *stream) {
    LV(kv_cg_p<float>, n, z, p, rz_new, rz_old, first, n);
}
int ust_k_cg_update_xr_f64(double *x, double *r, const double *p, const double *q, const double *num,
                           const double *den, size_t n, void *stream) {
    LV(kv_cg_xr<double>, n, x, r, p, q, num, den, n);
}
int ust_k_cg_update_xr_f32(float *x, float *r, const float *p, const float *q, const double *num, const double *den,
                           size_t n, void *stream) {
    LV(kv_cg_xr<float>, n, x, r, p, q, num, den, n);
}

int ust_k_cg_start(const double *bb, double rtol2, double *ctl, double *rz_old, void *stream) {
    hipLaunchKernelGGL(kv_cg_start, dim3(1), dim3(1), 0, ust_stream(stream), bb, rtol2, ctl, rz_old);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_cg_beta(const double *rz, double *rz_old, double *ctl, void *stream) {
    hipLaunchKernelGGL(kv_cg_beta, dim3(1), dim3(1), 0, ust_stream(stream), rz, rz_old, ctl);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_cg_p_guarded_f64(const double *z, double *p, const double *ctl, size_t n, void *stream) {
    LV(kv_cg_p_guarded, n, z, p, ctl, n);
}
int ust_k_cg_xr_guarded_f64(double *x, double *r, const double *p, const double *q, const double *rz,
                            const double *pq, const double *ctl, size_t n, void *stream) {
    LV(kv_cg_xr_guarded, n, x, r, p, q, rz, pq, ctl, n);
}
int ust_k_dot_mixed_beta(const double *x, const float *y, size_t n, double *partial, double *rz, double *rz_old,
                         double *ctl, void *stream) {
    const unsigned g = rgrid(n);
    hipLaunchKernelGGL(kv_dot_mixed_stage1, dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    hipLaunchKernelGGL(kv_reduce_beta, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, rz, rz_old, ctl);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_dot_beta_f64(const double *x, const double *y, size_t n, double *partial, double *rz, double *rz_old,
                       double *ctl, void *stream) {
    const unsigned g = rgrid(n);
    hipLaunchKernelGGL((kv_dot_stage1<double, true>), dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    hipLaunchKernelGGL(kv_reduce_beta, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, rz, rz_old, ctl);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_dot_mixed(const double *x, const float *y, size_t n, double *partial, double *out, void *stream) {
    const unsigned g = rgrid(n);
    hipLaunchKernelGGL(kv_dot_mixed_stage1, dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    hipLaunchKernelGGL(kv_reduce_stage2, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, out);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_cg_p_guarded_zf(const float *z, double *p, const double *ctl, size_t n, void *stream) {
    LV(kv_cg_p_guarded_zf, n, z, p, ctl, n);
}
int ust_k_cg_xr_rr_guarded_f64(double *x, double *r, const double *p, const double *q, const double *rz,
                               const double *pq, double *ctl, size_t n, double *partial, double *rr,
                               void *stream) {
    const unsigned g = rgrid(n);
    hipLaunchKernelGGL(kv_cg_xr_rr_guarded, dim3(g), dim3(256), 0, ust_stream(stream), x, r, p, q, rz, pq, ctl, n,
                       partial);
    hipLaunchKernelGGL(kv_reduce_check, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, rr, ctl);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_cheby_first_from_f64(const double *b, float *b32, const float *invd, float *d, float *x, float c2,
                               size_t n, void *stream) {
    LV(kv_cheby_first_from64, n, b, b32, invd, d, x, c2, n);
}
int ust_k_cg_p_guarded_f32(const float *z, float *p, const double *ctl, size_t n, void *stream) {
    if (al16(z) && al16(p)) LV(kv4_cg_p_guarded_f32, (n + 3) / 4, z, p, ctl, n);
    LV(kv_cg_p_guarded_f32, n, z, p, ctl, n);
}
int ust_k_cg_xr_rr_guarded_f32(float *x, float *r, const float *p, const float *q, const double *rz, const double *pq,
                               double *ctl, size_t n, double *partial, double *rr, void *stream) {
    const unsigned g = rgrid(n);
    if (al16(x) && al16(r) && al16(p) && al16(q))
        hipLaunchKernelGGL(kv4_cg_xr_rr_guarded_f32, dim3(g), dim3(256), 0, ust_stream(stream), x, r, p, q, rz, pq, ctl,
                           n, partial);
    else
        hipLaunchKernelGGL(kv_cg_xr_rr_guarded_f32, dim3(g), dim3(256), 0, ust_stream(stream), x, r, p, q, rz, pq, ctl,
                           n, partial);
    hipLaunchKernelGGL(kv_reduce_check, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, rr, ctl);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_dot_beta_f32(const float *x, const float *y, size_t n, double *partial, double *rz, double *rz_old,
                       double *ctl, void *stream) {
    const unsigned g = rgrid(n);
    if (al16(x) && al16(y))
        hipLaunchKernelGGL(kv4_dot_f32, dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    else
        hipLaunchKernelGGL((kv_dot_stage1<float, true>), dim3(g), dim3(256), 0, ust_stream(stream), x, y, n, partial);
    hipLaunchKernelGGL(kv_reduce_beta, dim3(1), dim3(256), 0, ust_stream(stream), partial, (int)g, rz, rz_old, ctl);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
int ust_k_copy_f64_to_f32_scaled(const double *x, float *y, double s, size_t n, void *stream) {
    LV(kv_d2f_scaled, n, x, y, s, n);
}
int ust_k_axpy_f32_to_f64(double s, const float *x, double *y, size_t n, void *stream) {
    LV(kv_axpy_f32_to_f64, n, s, x, y, n);
}
int ust_k_cg_check(const double *rr, double *ctl, void *stream) {
    hipLaunchKernelGGL(kv_cg_check, dim3(1), dim3(1), 0, ust_stream(stream), rr, ctl);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_cheby_step_f64(const double *b, const double *Ax, const double *invd, double *d, double *x, double c1,
                         double c2, size_t n, void *stream) {
    LV(kv_cheby<double>, n, b, Ax, invd, d, x, c1, c2, c1 == 0.0 ? 1 : 0, n);
}
int ust_k_cheby_step_f32(const float *b, const float *Ax, const float *invd, float *d, float *x, float c1, float c2,
                         size_t n, void *stream) {
    LV(kv_cheby<float>, n, b, Ax, invd, d, x, c1, c2, c1 == 0.0f ? 1 : 0, n);
}
int ust_k_cheby_first_f64(const double *b, const double *invd, double *d, double *x, double c2, size_t n,
                          void *stream) {
    LV(kv_cheby_first<double>, n, b, invd, d, x, c2, n);
}
int ust_k_cheby_first_f32(const float *b, const float *invd, float *d, float *x, float c2, size_t n, void *stream) {
    LV(kv_cheby_first<float>, n, b, invd, d, x, c2, n);
}
int ust_k_residual_f64(const double *b, const double *Ax, double *r, size_t n, void *stream) {
    LV(kv_residual<double>, n, b, Ax, r, n);
}
int ust_k_residual_f32(const float *b, const float *Ax, float *r, size_t n, void *stream) {
    LV(kv_residual<float>, n, b, Ax, r, n);
}

}  // extern "C"
