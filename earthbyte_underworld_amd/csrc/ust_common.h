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
// The K weights are stored tile-blocked, W[(tile*NQ + q)*256 + local], so that every load of a
// quadrature weight by a wavefront is one contiguous 64-element run.
// 3D tiles are 16 x 4 x 4 elements: one z layer (64 elements) per wavefront; the long x extent keeps
// the tile faces normal to x, whose nodes are gathered with a stride in memory, few.
#ifndef UST_TX3
#define UST_TX3 16
#define UST_TY3 4
#endif
#define UST_TZ3 4
#define UST_TX2 16
#define UST_TY2 16
#define UST_TILE 256

struct GridDev {
    int dim;
    int nx, ny, nz;       // elements per axis (nz = 1 in 2D)
    int nnx, nny, nnz;    // nodes per axis
    int nel;
    int tnx, tny, tnz;    // tiles per axis
    int ntiles;
    double hx, hy, hz;
    double xmin, ymin, zmin;
    double ymax;
};

// index of quadrature weight q of element e in the tile-blocked layout
__host__ __device__ __forceinline__ size_t widx(const GridDev &g, int e, int q) {
    const int ex = e % g.nx, ey = (e / g.nx) % g.ny, ez = e / (g.nx * g.ny);
    if (g.dim == 3) {
        const int t = ex / UST_TX3 + g.tnx * (ey / UST_TY3 + g.tny * (ez / UST_TZ3));
        const int l = ex % UST_TX3 + UST_TX3 * (ey % UST_TY3 + UST_TY3 * (ez % UST_TZ3));
        return ((size_t)t * 27 + q) * UST_TILE + l;
    }
    const int t = ex / UST_TX2 + g.tnx * (ey / UST_TY2);
    const int l = ex % UST_TX2 + UST_TX2 * (ey % UST_TY2);
    return ((size_t)t * 9 + q) * UST_TILE + l;
}

static inline GridDev grid_dev(const ust_grid *g) {
    GridDev d;
    d.dim = g->dim;
    d.nx = g->nel[0];
    d.ny = g->nel[1];
    d.nz = g->dim == 3 ? g->nel[2] : 1;
    d.nnx = d.nx + 1;
    d.nny = d.ny + 1;
    d.nnz = g->dim == 3 ? d.nz + 1 : 1;
    d.nel = d.nx * d.ny * d.nz;
    if (g->dim == 3) {
        d.tnx = (d.nx + UST_TX3 - 1) / UST_TX3;
        d.tny = (d.ny + UST_TY3 - 1) / UST_TY3;
        d.tnz = (d.nz + UST_TZ3 - 1) / UST_TZ3;
    } else {
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
