// Device-side helpers shared by the gfx950 kernels of the PIC Stokes hot path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ust_kernels.h"

#define UST_BLOCK 256

#define UST_LAUNCH_CHECK()                                   \
    do {                                                     \
        hipError_t _e = hipGetLastError();                   \
        if (_e != hipSuccess) {                              \
            ust_set_hip_error(_e, __FILE__, __LINE__);       \
            return UST_EDEVICE;                              \
        }                                                    \
    } while (0)

void ust_set_hip_error(hipError_t e, const char *file, int line);
// an error message for ust_dev_error_string() from a check that is not a HIP call
void ust_set_error_text(const char *text);

static inline unsigned ust_grid_size(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g == 0 ? 1 : g);
}

static inline hipStream_t ust_stream(void *s) { return (hipStream_t)s; }

// Quadratic Lagrange basis on the points {-1, 0, 1}.
__device__ __forceinline__ void lagrange3(double x, double *L) {
    L[0] = 0.5 * x * (x - 1.0);
    L[1] = 1.0 - x * x;
    L[2] = 0.5 * x * (x + 1.0);
}

// c(a, q) = 1 + s_a * xi_q for xi_q = q - 1 and s_a = 2a - 1: the factor of the Q1 shape
// function of vertex bit a evaluated at Lagrange point q (values 0, 1, 2).
__host__ __device__ constexpr int cfac(int a, int q) { return a ? q : 2 - q; }
__host__ __device__ constexpr int sgn(int a) { return 2 * a - 1; }

// Element tiles of the operator kernels: one 256-thread workgroup per tile, one element per thread.
// 2D tiles are 16 x 16 elements.  In 3D a workgroup owns a 16 x 16 column of elements and streams a chunk
// of tz element layers along z (tz from the grid size, so that a launch has enough workgroups); the K
// weights of a column are stored layer after layer, W[((txy*nz + ez)*27 + q)*256 + local], so that every
// load of a quadrature weight by a wavefront is one contiguous 64-element run, for any chunk length.  A 3D
// grid with fewer than UST_COL_MIN column layers (tnx*tny*nz) uses 16 x 4 x 4 tiles instead, stored tile
// after tile, W[(tile*27 + q)*256 + local]: small grids then get more workgroups.
#define UST_TX3 16
#define UST_TY3 16
#define UST_TTX3 16
#define UST_TTY3 4
#define UST_TTZ3 4
#define UST_TX2 16
#define UST_TY2 16
#define UST_TILE 256
#define UST_TZ3_MAX 16
#define UST_COL_MIN 1024
// workgroups a column launch aims at (the z-chunk length follows): two per CU; 1024 (four per CU, shorter
// chunks) measured slower on the 128x128x64 solve (FP64 product 93 -> 100 us).  The FP32 product and residual
// passes, which fit four workgroups per CU, aim at 1024 (UST_COL_WG_TARGET_F32_YR).
#ifndef UST_COL_WG_TARGET
#define UST_COL_WG_TARGET 512
#endif
#ifndef UST_COL_WG_TARGET_F32_YR
#define UST_COL_WG_TARGET_F32_YR 1024
#endif

struct GridDev {
    int dim;
    int col;              // 3D column layout (1) or 16x4x4 tiles (0)
    int nx, ny, nz;       // elements per axis (nz = 1 in 2D)
    int nnx, nny, nnz;    // nodes per axis
    int nel;
    int tx, ty, tz;       // tile extent in elements (3D: tz = z-chunk length of a workgroup)
    int tnx, tny, tnz;    // tiles per axis (3D: tnz = z-chunks per column)
    int ntiles;
    double hx, hy, hz;
    double xmin, ymin, zmin;
    double ymax;
};

// index of quadrature weight q of element e in the tile-blocked layout
__host__ __device__ __forceinline__ size_t widx(const GridDev &g, int e, int q) {
    const int ex = e % g.nx, ey = (e / g.nx) % g.ny, ez = e / (g.nx * g.ny);
    if (g.dim == 3 && g.col) {
        const int txy = ex / UST_TX3 + g.tnx * (ey / UST_TY3);
        const int l = ex % UST_TX3 + UST_TX3 * (ey % UST_TY3);
        return (((size_t)txy * g.nz + ez) * 27 + q) * UST_TILE + l;
    }
    if (g.dim == 3) {
        const int t = ex / UST_TTX3 + g.tnx * (ey / UST_TTY3 + g.tny * (ez / UST_TTZ3));
        const int l = ex % UST_TTX3 + UST_TTX3 * (ey % UST_TTY3 + UST_TTY3 * (ez % UST_TTZ3));
        return ((size_t)t * 27 + q) * UST_TILE + l;
    }
    const int t = ex / UST_TX2 + g.tnx * (ey / UST_TY2);
    const int l = ex % UST_TX2 + UST_TX2 * (ey % UST_TY2);
    return ((size_t)t * 9 + q) * UST_TILE + l;
}

// wg_target: the workgroups a column launch aims at (the z-chunk length follows)
static inline GridDev grid_dev(const ust_grid *g, long long wg_target = UST_COL_WG_TARGET) {
    GridDev d;
    d.dim = g->dim;
    d.nx = g->nel[0];
    d.ny = g->nel[1];
    d.nz = g->dim == 3 ? g->nel[2] : 1;
    d.nnx = d.nx + 1;
    d.nny = d.ny + 1;
    d.nnz = g->dim == 3 ? d.nz + 1 : 1;
    d.nel = d.nx * d.ny * d.nz;
    d.col = 0;
    if (g->dim == 3) {
        const int cnx = (d.nx + UST_TX3 - 1) / UST_TX3, cny = (d.ny + UST_TY3 - 1) / UST_TY3;
        const long long layers = (long long)cnx * cny * d.nz;
        d.col = layers >= UST_COL_MIN;
        if (d.col) {
            d.tx = UST_TX3;
            d.ty = UST_TY3;
            d.tnx = cnx;
            d.tny = cny;
            // z-chunk: long enough that few node planes sit on chunk boundaries, short enough that the launch
            // has about wg_target workgroups or more
            long long zc = layers / wg_target;
            if (zc < 2) zc = 2;
            if (zc > UST_TZ3_MAX) zc = UST_TZ3_MAX;
            if (zc > d.nz) zc = d.nz;
            d.tz = (int)zc;
            d.tnz = (d.nz + d.tz - 1) / d.tz;
        } else {
            d.tx = UST_TTX3;
            d.ty = UST_TTY3;
            d.tz = UST_TTZ3;
            d.tnx = (d.nx + UST_TTX3 - 1) / UST_TTX3;
            d.tny = (d.ny + UST_TTY3 - 1) / UST_TTY3;
            d.tnz = (d.nz + UST_TTZ3 - 1) / UST_TTZ3;
        }
    } else {
        d.tx = UST_TX2;
        d.ty = UST_TY2;
        d.tz = 1;
        d.tnx = (d.nx + UST_TX2 - 1) / UST_TX2;
        d.tny = (d.ny + UST_TY2 - 1) / UST_TY2;
        d.tnz = 1;
    }
    d.ntiles = d.tnx * d.tny * d.tnz;
    d.hx = g->h[0];
    d.hy = g->h[1];
    d.hz = g->dim == 3 ? g->h[2] : 1.0;
    d.xmin = g->xmin[0];
    d.ymin = g->xmin[1];
    d.zmin = g->dim == 3 ? g->xmin[2] : 0.0;
    d.ymax = g->ymax;
    return d;
}

// bfloat16 is the upper half of an FP32: widened by a shift, narrowed with round-to-nearest-even
__device__ __forceinline__ float bf16_to_f32(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }
__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
// K weights stored as bfloat16 are read through wload<float, unsigned short>
template <typename T, typename WT>
__device__ __forceinline__ T wload(const WT *__restrict__ W, size_t i) {
    return (T)W[i];
}
template <>
__device__ __forceinline__ float wload<float, unsigned short>(const unsigned short *__restrict__ W, size_t i) {
    return bf16_to_f32(W[i]);
}
