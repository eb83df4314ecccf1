// Swarm advection and cell re-sort on gfx950: the step after the Stokes solve.
//
// SwarmAdvector (PICellerator/MaterialPoints/src/SwarmAdvector.c:288-313) with the TimeIntegrand's first- and
// second-order updates (StgDomain/Utils/src/TimeIntegrand.c:221-330): the velocity at a particle is the Q1
// interpolation of the nodal velocity in the element holding the point (FieldVariable_InterpolateValueAt);
// a point outside the box has no velocity (OUTSIDE_GLOBAL).  Particles that end outside the box are removed,
// as ParticleMovementHandler does with particles that leave the global domain.  The swarm is then re-sorted by
// element, stably (particles keep their relative order inside a cell), local coordinates are recomputed in the
// new element, and ConstantWeights gives every particle of a cell 2^dim / count
// (PICellerator/Weights/src/ConstantWeights.c:187, WeightsCalculator.c:83).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "ust_common.h"

namespace {

__device__ __forceinline__ int find_cell_adv(const int *cell_off, int nel, int p) {
    int lo = 0, hi = nel - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cell_off[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct Box {
    double xmin[3], xmax[3];
    int n[3];
};

// vertex coordinate as CartesianGenerator computes it
__device__ __forceinline__ double vcoord_adv(const Box &b, int d, int i) {
    return b.xmin[d] + (double)i * (b.xmax[d] - b.xmin[d]) / (double)b.n[d];
}

// element index along axis d holding coordinate x (x inside the box), and the local coordinate there
__device__ __forceinline__ int locate(const Box &b, int d, double x, double *xi) {
    int i = (int)floor((x - b.xmin[d]) * (double)b.n[d] / (b.xmax[d] - b.xmin[d]));
    i = i < 0 ? 0 : (i >= b.n[d] ? b.n[d] - 1 : i);
    const double mn = vcoord_adv(b, d, i), mx = vcoord_adv(b, d, i + 1);
    const double c = 0.5 * (mn + mx), hh = 0.5 * (mx - mn);
    double t = (x - c) / hh;
    t += (x - (c + hh * t)) / hh;  // one Newton correction of the affine map, as the layout kernel
    *xi = t;
    return i;
}

template <int DIM>
__device__ __forceinline__ bool inside(const Box &b, const double *x) {
    bool ok = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) ok = ok && x[d] >= b.xmin[d] && x[d] <= b.xmax[d];
    return ok;
}

// Q1 interpolation of the nodal velocity u (interleaved n*DIM + c) at x; false outside the box
template <int DIM>
__device__ bool velocity_at(const Box &b, const double *__restrict__ u, const double *x, double *v) {
    if (!inside<DIM>(b, x)) return false;
    int idx[3] = {0, 0, 0};
    double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; d++) idx[d] = locate(b, d, x[d], &xi[d]);
    const int nnx = b.n[0] + 1, nny = b.n[1] + 1;
#pragma unroll
    for (int c = 0; c < DIM; c++) v[c] = 0.0;
#pragma unroll
    for (int a = 0; a < (1 << DIM); a++) {
        double N = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) N *= ((a >> d) & 1) ? 0.5 * (1.0 + xi[d]) : 0.5 * (1.0 - xi[d]);
        const size_t node = (size_t)(idx[0] + (a & 1)) +
                            (size_t)nnx * ((size_t)(idx[1] + ((a >> 1) & 1)) +
                                           (DIM == 3 ? (size_t)nny * (size_t)(idx[2] + ((a >> 2) & 1)) : 0));
#pragma unroll
        for (int c = 0; c < DIM; c++) v[c] += N * u[node * DIM + c];
    }
    return true;
}

// One thread per particle: new position by Euler (order 1) or the midpoint rule (order 2), then its element
// (key; nel = removed) and local coordinates there.  flag bit 0: a corrector had no velocity and the first-
// order fallback is not allowed (the particle is left where it was).
template <int DIM>
__global__ void k_advect(Box b, int nel, const int *__restrict__ cell_off, int np, const double *__restrict__ xi0,
                         const double *__restrict__ xi1, const double *__restrict__ xi2,
                         const double *__restrict__ u, double dt, int order, int fallback,
                         unsigned *__restrict__ key, double *__restrict__ n0, double *__restrict__ n1,
                         double *__restrict__ n2, int *__restrict__ flag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int e = find_cell_adv(cell_off, nel, p);
    const int idx[3] = {e % b.n[0], (e / b.n[0]) % b.n[1], e / (b.n[0] * b.n[1])};
    const double xi[3] = {xi0[p], xi1[p], DIM == 3 ? xi2[p] : 0.0};
    double x0[3] = {0.0, 0.0, 0.0}, v0[3] = {0.0, 0.0, 0.0}, x1[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double mn = vcoord_adv(b, d, idx[d]), mx = vcoord_adv(b, d, idx[d] + 1);
        x0[d] = 0.5 * (mn + mx) + 0.5 * (mx - mn) * xi[d];
    }
    velocity_at<DIM>(b, u, x0, v0);  // x0 lies in its element
    if (order == 1) {
#pragma unroll
        for (int d = 0; d < DIM; d++) x1[d] = x0[d] + dt * v0[d];
    } else {
        double xm[3] = {0.0, 0.0, 0.0}, vm[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < DIM; d++) xm[d] = x0[d] + 0.5 * dt * v0[d];
        if (velocity_at<DIM>(b, u, xm, vm)) {
#pragma unroll
            for (int d = 0; d < DIM; d++) x1[d] = x0[d] + dt * vm[d];
        } else if (fallback) {  // _TimeIntegrand_RewindToStartAndApplyFirstOrderUpdate
#pragma unroll
            for (int d = 0; d < DIM; d++) x1[d] = x0[d] + dt * v0[d];
        } else {
            atomicOr(flag, 1);
#pragma unroll
            for (int d = 0; d < DIM; d++) x1[d] = x0[d];
        }
    }
    double nxi[3] = {0.0, 0.0, 0.0};
    unsigned k = (unsigned)nel;
    if (inside<DIM>(b, x1)) {
        int id[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; d++) id[d] = locate(b, d, x1[d], &nxi[d]);
        k = (unsigned)(id[0] + b.n[0] * (id[1] + b.n[1] * (DIM == 3 ? id[2] : 0)));
    }
    key[p] = k;
    n0[p] = nxi[0];
    n1[p] = nxi[1];
    if (DIM == 3) n2[p] = nxi[2];
}

__global__ void k_iota(unsigned *v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (unsigned)i;
}

// cell_off from the sorted keys: thread s writes the offsets of the cells (key[s-1], key[s]] (empty cells
// included), the last thread those after its key.  Keys equal to nel (removed particles) end the array.
__global__ void k_cell_offsets(const unsigned *__restrict__ key, int np, int nel, int *__restrict__ cell_off) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > np) return;
    const int lo = s == 0 ? 0 : (int)key[s - 1] + 1;
    const int hi = s == np ? nel : (int)key[s];
    for (int c = lo; c <= hi && c <= nel; c++) cell_off[c] = s;
}

template <int DIM>
__global__ void k_swarm_gather(int np, const unsigned *__restrict__ perm, const unsigned *__restrict__ key,
                               const int *__restrict__ cell_off, const double *__restrict__ n0,
                               const double *__restrict__ n1, const double *__restrict__ n2,
                               const unsigned char *__restrict__ mat, const float *__restrict__ coh,
                               double *__restrict__ o0, double *__restrict__ o1, double *__restrict__ o2,
                               double *__restrict__ w, unsigned char *__restrict__ omat, float *__restrict__ ocoh) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    const unsigned src = perm[s];
    const int c = (int)key[s];
    o0[s] = n0[src];
    o1[s] = n1[src];
    if (DIM == 3) o2[s] = n2[src];
    w[s] = (double)(1 << DIM) / (double)(cell_off[c + 1] - cell_off[c]);
    if (mat) omat[s] = mat[src];
    if (coh) ocoh[s] = coh[src];
}

static Box box_of(const ust_grid *g) {
    Box b;
    for (int d = 0; d < 3; d++) {
        b.n[d] = d < g->dim ? g->nel[d] : 1;
        b.xmin[d] = g->xmin[d];
        b.xmax[d] = d < g->dim ? g->xmax[d] : g->xmin[d] + 1.0;
    }
    return b;
}

static unsigned key_bits(int nel) {
    unsigned bits = 1;
    while (bits < 32 && ((unsigned long long)1 << bits) <= (unsigned long long)nel) bits++;
    return bits;
}

}  // namespace

extern "C" {

int ust_k_advect(const ust_grid *g, const int *cell_off, int np, const double *xi0, const double *xi1,
                 const double *xi2, const double *u, double dt, int order, int fallback, unsigned *key, double *n0,
                 double *n1, double *n2, int *flag, void *stream) {
    if (!g || !cell_off || !xi0 || !xi1 || !u || !key || !n0 || !n1 || !flag || np < 0) return UST_EINVAL;
    if (g->dim == 3 && (!xi2 || !n2)) return UST_EINVAL;
    if (order != 1 && order != 2) return UST_EINVAL;
    if (np == 0) return UST_OK;
    const Box b = box_of(g);
    const int nel = b.n[0] * b.n[1] * b.n[2];
    hipStream_t st = ust_stream(stream);
    if (g->dim == 3)
        hipLaunchKernelGGL(k_advect<3>, dim3(ust_grid_size(np, 256)), dim3(256), 0, st, b, nel, cell_off, np, xi0,
                           xi1, xi2, u, dt, order, fallback, key, n0, n1, n2, flag);
    else
        hipLaunchKernelGGL(k_advect<2>, dim3(ust_grid_size(np, 256)), dim3(256), 0, st, b, nel, cell_off, np, xi0,
                           xi1, xi2, u, dt, order, fallback, key, n0, n1, n2, flag);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

size_t ust_k_sort_by_cell_work_bytes(int np, int nel) {
    size_t bytes = 0;
    if (np <= 0) return 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, (unsigned *)nullptr, (unsigned *)nullptr, (unsigned *)nullptr,
                                  (unsigned *)nullptr, (unsigned)np, 0, key_bits(nel)) != hipSuccess)
        return 0;
    return bytes;
}

int ust_k_sort_by_cell(const unsigned *key, unsigned *key_sorted, unsigned *perm, unsigned *perm_tmp, int np, int nel,
                       void *work, size_t work_bytes, void *stream) {
    if (np < 0 || nel <= 0 || (np > 0 && (!key || !key_sorted || !perm || !perm_tmp || !work))) return UST_EINVAL;
    if (np == 0) return UST_OK;
    hipStream_t st = ust_stream(stream);
    hipLaunchKernelGGL(k_iota, dim3(ust_grid_size(np, 256)), dim3(256), 0, st, perm_tmp, np);
    size_t bytes = work_bytes;
    if (rocprim::radix_sort_pairs(work, bytes, key, key_sorted, (const unsigned *)perm_tmp, perm, (unsigned)np, 0,
                                  key_bits(nel), st) != hipSuccess)
        return UST_EDEVICE;
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_cell_offsets(const unsigned *key_sorted, int np, int nel, int *cell_off, void *stream) {
    if (np < 0 || nel <= 0 || !cell_off || (np > 0 && !key_sorted)) return UST_EINVAL;
    hipLaunchKernelGGL(k_cell_offsets, dim3(ust_grid_size((size_t)np + 1, 256)), dim3(256), 0, ust_stream(stream),
                       key_sorted, np, nel, cell_off);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_swarm_gather(int dim, int np, const unsigned *perm, const unsigned *key_sorted, const int *cell_off,
                       const double *n0, const double *n1, const double *n2, const unsigned char *mat,
                       const float *coh, double *o0, double *o1, double *o2, double *w, unsigned char *omat,
                       float *ocoh, void *stream) {
    if (np < 0 || (np > 0 && (!perm || !key_sorted || !cell_off || !n0 || !n1 || !o0 || !o1 || !w))) return UST_EINVAL;
    if (dim == 3 && np > 0 && (!n2 || !o2)) return UST_EINVAL;
    if ((mat && !omat) || (coh && !ocoh)) return UST_EINVAL;
    if (np == 0) return UST_OK;
    hipStream_t st = ust_stream(stream);
    if (dim == 3)
        hipLaunchKernelGGL(k_swarm_gather<3>, dim3(ust_grid_size(np, 256)), dim3(256), 0, st, np, perm, key_sorted,
                           cell_off, n0, n1, n2, mat, coh, o0, o1, o2, w, omat, ocoh);
    else
        hipLaunchKernelGGL(k_swarm_gather<2>, dim3(ust_grid_size(np, 256)), dim3(256), 0, st, np, perm, key_sorted,
                           cell_off, n0, n1, n2, mat, coh, o0, o1, o2, w, omat, ocoh);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

}  // extern "C"
