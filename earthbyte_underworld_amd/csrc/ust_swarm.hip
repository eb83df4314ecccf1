// Swarm set-up on gfx950: random particle layout with MINSTD jump-ahead, the coincident
// global->local map and constant weights, material shapes, per-particle noise.
#include <math.h>

#include "ust_common.h"

// a^k mod m for the Park-Miller generator (m = 2^31 - 1; products fit in 64 bits)
__device__ __forceinline__ unsigned long long powmod(unsigned long long a, unsigned long long k,
                                                     unsigned long long m) {
    unsigned long long r = 1;
    a %= m;
    while (k) {
        if (k & 1) r = (r * a) % m;
        a = (a * a) % m;
        k >>= 1;
    }
    return r;
}

__device__ __forceinline__ int find_cell(const int *cell_off, int nel, int p) {
    int lo = 0, hi = nel - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cell_off[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct BoxDev {
    double xmin[3], xmax[3];
};

// vertex coordinate as CartesianGenerator and the oracle compute it
__device__ __forceinline__ double vcoord(double mn, double mx, int i, int n) {
    return mn + (double)i * (mx - mn) / (double)n;
}

template <int DIM>
__global__ void k_random_layout(GridDev g, BoxDev box, const int *__restrict__ cell_off, int np, long long s0,
                                double *__restrict__ xi0, double *__restrict__ xi1, double *__restrict__ xi2,
                                double *__restrict__ w) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const unsigned long long M = 2147483647ULL, A = 16807ULL;
    const int e = find_cell(cell_off, g.nel, p);
    const int idx[3] = {e % g.nx, (e / g.nx) % g.ny, e / (g.nx * g.ny)};
    const int nn[3] = {g.nx, g.ny, g.nz};
    unsigned long long s = (powmod(A, 3ULL * (unsigned long long)p + 1ULL, M) * ((unsigned long long)s0 % M)) % M;
    double xi[3];
    for (int d = 0; d < 3; d++) {
        if (d > 0) s = (s * A) % M;
        if (d < DIM) {
            const double mn = vcoord(box.xmin[d], box.xmax[d], idx[d], nn[d]);
            const double mx = vcoord(box.xmin[d], box.xmax[d], idx[d] + 1, nn[d]);
            const double x = (mx - mn) * (double)s / 2147483647.0 + mn;  // Random.c:71-73
            const double c = 0.5 * (mn + mx), hh = 0.5 * (mx - mn);
            double t = (x - c) / hh;
            // one Newton correction of the affine map (ElementType.c:227-350 converges in one step)
            t += (x - (c + hh * t)) / hh;
            xi[d] = t;
        }
    }
    xi0[p] = xi[0];
    xi1[p] = xi[1];
    if (DIM == 3) xi2[p] = xi[2];
    w[p] = (double)(1 << DIM) / (double)(cell_off[e + 1] - cell_off[e]);
}

// SpaceFillerParticleLayout (StgDomain/Swarm/src/SpaceFillerParticleLayout.c:218-240 with GlobalParticleLayout.c:
// 275-318): global particle p sits at min + (max - min) s_d per axis, s_d its (p+1)-th Sobol number over 2^30 --
// the XOR of the axis' direction numbers m_k << (29 - k) over the set bits k of the Gray code of p + 1, which
// is what SobolGenerator_GetNextNumber's recurrence yields (SobolGenerator.c:295-338).  Its element follows
// Mesh_RegularAlgorithms_SearchElements (Mesh_RegularAlgorithms.c:194-226): per axis the integer part of
// (x - min) / sep, one lower when the fraction is within 1e-12 of zero (a point on an interior face belongs to
// the lower element), the last element within 1e-12 of max.  The coordinate products are not contracted into
// fused multiply-adds, so that x and the element decision round as the reference's host code does.
struct SobolDirs {
    unsigned m[3][32];
};

template <int DIM>
__global__ void k_spacefiller(GridDev g, BoxDev box, int np, SobolDirs dirs, unsigned *__restrict__ key,
                              double *__restrict__ xi0, double *__restrict__ xi1, double *__restrict__ xi2) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const unsigned n = (unsigned)p + 1u, gray = n ^ (n >> 1);
    const int nn[3] = {g.nx, g.ny, g.nz};
    int ind[3] = {0, 0, 0};
    double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        unsigned s = 0;
        for (int k = 0; k < 30; k++)
            if ((gray >> k) & 1u) s ^= dirs.m[d][k] << (29 - k);
        const double lo = box.xmin[d], hi = box.xmax[d];
        const double x = (hi - lo) * ((double)s / 1073741824.0) + lo;
        int i;
        if (x - hi >= -1e-12 && x - hi <= 1e-12) {
            i = nn[d] - 1;
        } else {
            const double out = (x - lo) / ((hi - lo) / (double)nn[d]);
            const double ip = floor(out), frac = out - ip;
            i = (int)ip;
            if (i > 0 && frac >= -1e-12 && frac <= 1e-12) i--;
            i = i < 0 ? 0 : (i >= nn[d] ? nn[d] - 1 : i);
        }
        ind[d] = i;
        const double mn = vcoord(lo, hi, i, nn[d]), mx = vcoord(lo, hi, i + 1, nn[d]);
        const double c = 0.5 * (mn + mx), hh = 0.5 * (mx - mn);
        double t = (x - c) / hh;
        t += (x - (c + hh * t)) / hh;  // one Newton correction of the affine map, as the random layout
        xi[d] = t;
    }
    key[p] = (unsigned)(ind[0] + g.nx * (ind[1] + (DIM == 3 ? g.ny * ind[2] : 0)));
    xi0[p] = xi[0];
    xi1[p] = xi[1];
    if (DIM == 3) xi2[p] = xi[2];
}

template <int DIM>
__global__ void k_assign_materials(GridDev g, BoxDev box, const int *__restrict__ cell_off, int np,
                                   const double *__restrict__ xi0, const double *__restrict__ xi1,
                                   const double *__restrict__ xi2, const ust_shape *__restrict__ shapes, int nshape,
                                   const unsigned char *__restrict__ cell_mat, unsigned char *__restrict__ mat) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int e = find_cell(cell_off, g.nel, p);
    const int idx[3] = {e % g.nx, (e / g.nx) % g.ny, e / (g.nx * g.ny)};
    const int nn[3] = {g.nx, g.ny, g.nz};
    const double xi[3] = {xi0[p], xi1[p], DIM == 3 ? xi2[p] : 0.0};
    double x[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < DIM; d++) {
        const double mn = vcoord(box.xmin[d], box.xmax[d], idx[d], nn[d]);
        const double mx = vcoord(box.xmin[d], box.xmax[d], idx[d] + 1, nn[d]);
        x[d] = mn + 0.5 * (xi[d] + 1.0) * (mx - mn);
    }
    int m = cell_mat ? (int)cell_mat[e] : 0;
    for (int s = 0; s < nshape; s++) {
        const ust_shape &sh = shapes[s];
        bool in = false;
        if (sh.type == UST_SHAPE_EVERYWHERE) {
            in = true;
        } else if (sh.type == UST_SHAPE_BELOWCOSINEPLANE) {
            const double y = sh.p[0] + sh.p[1] * cos((2 * M_PI * x[0] / sh.p[2]) + sh.p[3]);
            in = x[1] <= y;
        } else if (sh.type == UST_SHAPE_SPHERE) {
            double r2 = 0.0;
            for (int d = 0; d < DIM; d++) r2 += (x[d] - sh.p[d]) * (x[d] - sh.p[d]);
            in = r2 <= sh.p[3] * sh.p[3];
        } else if (sh.type == UST_SHAPE_BOX) {
            in = true;
            for (int d = 0; d < DIM; d++) in = in && x[d] >= sh.p[d] && x[d] <= sh.p[3 + d];
        }
        if (in) m = sh.material;
    }
    mat[p] = (unsigned char)m;
}

// material of each particle from a nodal field sampled at the particle (Q1 interpolation in its element),
// quantised to nbins materials first_mat.. over [lo, hi]
template <int DIM>
__global__ void k_assign_materials_field(GridDev g, const int *__restrict__ cell_off, int np,
                                         const double *__restrict__ xi0, const double *__restrict__ xi1,
                                         const double *__restrict__ xi2, const float *__restrict__ field, double lo,
                                         double hi, int nbins, int first_mat, unsigned char *__restrict__ mat) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int e = find_cell(cell_off, g.nel, p);
    const int ix = e % g.nx, iy = (e / g.nx) % g.ny, iz = e / (g.nx * g.ny);
    const double xi[3] = {xi0[p], xi1[p], DIM == 3 ? xi2[p] : 0.0};
    double v = 0.0;
    for (int a = 0; a < (1 << DIM); a++) {
        const int bx = a & 1, by = (a >> 1) & 1, bz = (a >> 2) & 1;
        double N = 0.5 * (1.0 + (2 * bx - 1) * xi[0]) * 0.5 * (1.0 + (2 * by - 1) * xi[1]);
        if (DIM == 3) N *= 0.5 * (1.0 + (2 * bz - 1) * xi[2]);
        const size_t n = (size_t)(ix + bx) + (size_t)g.nnx * ((size_t)(iy + by) + (size_t)g.nny * (iz + bz));
        v += N * (double)field[n];
    }
    double t = rint((v - lo) / (hi - lo) * (double)(nbins - 1));
    t = t < 0.0 ? 0.0 : (t > (double)(nbins - 1) ? (double)(nbins - 1) : t);
    mat[p] = (unsigned char)(first_mat + (int)t);
}

__device__ __forceinline__ unsigned long long mix64s(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void k_noise_scale(float *coh, int np, float amp, unsigned long long seed) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double u = 2.0 * ((double)(mix64s(seed * 0x100000001B3ULL + (unsigned long long)p) >> 11) *
                            (1.0 / 9007199254740992.0)) - 1.0;
    coh[p] = 1.0f + amp * (float)u;
}

static BoxDev box_of(const ust_grid *g) {
    BoxDev b;
    for (int d = 0; d < 3; d++) {
        b.xmin[d] = g->xmin[d];
        b.xmax[d] = g->xmax[d];
    }
    return b;
}

extern "C" int ust_k_random_layout(const ust_grid *gr, const int *cell_off, int np, long seed, double *xi0,
                                   double *xi1, double *xi2, double *w, void *stream) {
    GridDev g = grid_dev(gr);
    if (!cell_off || !xi0 || !xi1 || !w || (g.dim == 3 && !xi2) || np < 0) return UST_EINVAL;
    if (np == 0) return UST_OK;
    const long long s0 = (long long)(seed ^ 123459876L);
    if (s0 <= 0 || s0 >= 2147483647LL) return UST_EINVAL;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_random_layout<3>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream), g,
                           box_of(gr), cell_off, np, s0, xi0, xi1, xi2, w);
    else
        hipLaunchKernelGGL(k_random_layout<2>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream), g,
                           box_of(gr), cell_off, np, s0, xi0, xi1, xi2, w);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_spacefiller_layout(const ust_grid *gr, int np, const unsigned *dirnum, unsigned *key,
                                        double *xi0, double *xi1, double *xi2, void *stream) {
    GridDev g = grid_dev(gr);
    if (!dirnum || !key || !xi0 || !xi1 || (g.dim == 3 && !xi2) || np < 0) return UST_EINVAL;
    if (np == 0) return UST_OK;
    SobolDirs dirs;
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < 32; k++) dirs.m[d][k] = k < 30 ? dirnum[d * 32 + k] : 0u;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_spacefiller<3>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream), g,
                           box_of(gr), np, dirs, key, xi0, xi1, xi2);
    else
        hipLaunchKernelGGL(k_spacefiller<2>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream), g,
                           box_of(gr), np, dirs, key, xi0, xi1, xi2);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_assign_materials(const ust_grid *gr, const int *cell_off, int np, const double *xi0,
                                      const double *xi1, const double *xi2, const ust_shape *shapes_dev, int nshape,
                                      const unsigned char *cell_material, unsigned char *mat, void *stream) {
    GridDev g = grid_dev(gr);
    if (!cell_off || !xi0 || !xi1 || !mat || (g.dim == 3 && !xi2) || np < 0 || nshape < 0 ||
        (nshape > 0 && !shapes_dev))
        return UST_EINVAL;
    if (np == 0) return UST_OK;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_assign_materials<3>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream), g,
                           box_of(gr), cell_off, np, xi0, xi1, xi2, shapes_dev, nshape, cell_material, mat);
    else
        hipLaunchKernelGGL(k_assign_materials<2>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream), g,
                           box_of(gr), cell_off, np, xi0, xi1, xi2, shapes_dev, nshape, cell_material, mat);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_noise_scale(float *coh, int np, float amp, unsigned long long seed, void *stream) {
    if (!coh || np < 0) return UST_EINVAL;
    if (np == 0) return UST_OK;
    hipLaunchKernelGGL(k_noise_scale, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream), coh, np, amp,
                       seed);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_assign_materials_field(const ust_grid *gr, const int *cell_off, int np, const double *xi0,
                                            const double *xi1, const double *xi2, const float *field_nodal, double lo,
                                            double hi, int nbins, int first_material, unsigned char *mat,
                                            void *stream) {
    GridDev g = grid_dev(gr);
    if (!cell_off || !xi0 || !xi1 || !mat || !field_nodal || (g.dim == 3 && !xi2) || np < 0 || nbins < 1 ||
        first_material < 0 || first_material + nbins > 256 || !(hi > lo))
        return UST_EINVAL;
    if (np == 0) return UST_OK;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_assign_materials_field<3>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream),
                           g, cell_off, np, xi0, xi1, xi2, field_nodal, lo, hi, nbins, first_material, mat);
    else
        hipLaunchKernelGGL(k_assign_materials_field<2>, dim3(ust_grid_size(np, 256)), dim3(256), 0, ust_stream(stream),
                           g, cell_off, np, xi0, xi1, xi2, field_nodal, lo, hi, nbins, first_material, mat);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
