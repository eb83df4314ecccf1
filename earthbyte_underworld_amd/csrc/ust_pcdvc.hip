// PCDVC population control (PICellerator/Utils/src/PCDVC.c, non-Inflow branch) on the cell-sorted swarm.
//
// The reference walks the elements one by one, splitting particles whose discrete Voronoi region is larger than
// upperT % of the element and deleting particles whose region is smaller than lowerT %, then rebuilds the
// element's Voronoi regions.  Every element's decision depends only on its own particles, so here one thread per
// element plans its new particle list (k_pcdvc_plan), the host turns the counts into offsets, one thread per
// element writes the new swarm (k_pcdvc_build), and the DVC kernel then runs once over the new swarm.
//
// List entries: s (original particle s of the element) or s | PCDVC_SPLIT (a copy of s at its region centroid).
// The element's list region in the scratch array starts at cell_off[e] + e * max_splits and holds at most
// n + max_splits entries.
#include "ust_common.h"

#define PCDVC_SPLIT (1 << 30)

template <int DIM>
__global__ void k_pcdvc_plan(GridDev g, const int *__restrict__ cell_off, const double *__restrict__ w,
                             const unsigned char *__restrict__ mat, ust_pcdvc_opts o, int *__restrict__ new_count,
                             int *__restrict__ list, int *__restrict__ flag) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int p0 = cell_off[e], n = cell_off[e + 1] - p0;
    int *L = list + p0 + (size_t)e * o.max_splits;
    const double vol = (double)(1 << DIM);
    const double max_w = o.upper_t * vol / 100.0, min_w = o.lower_t * vol / 100.0;  // PCDVC.c:1338-1339
    // interface element: not every particle has the first particle's material (:1160-1167)
    bool interface = false;
    const int m0 = mat && n > 0 ? (int)mat[p0] : 0;
    if (mat)
        for (int i = 1; i < n && !interface; i++) interface = (int)mat[p0 + i] != m0;
    // boundary element: first or last element layer along any axis (:1171-1176)
    const int ix = e % g.nx, iy = (e / g.nx) % g.ny, iz = e / (g.nx * g.ny);
    const bool boundary = ix == 0 || ix == g.nx - 1 || iy == 0 || iy == g.ny - 1 ||
                          (DIM == 3 && (iz == 0 || iz == g.nz - 1));
    int mdel = o.max_deletions, mspl = o.max_splits;
    if (interface) {  // :1341-1345
        if (!o.delete_in_interface_cells) mdel = 0;
        if (!o.split_in_interface_cells && !boundary) mspl = 0;
    }
    if (mdel > n - 4) mdel = n / 2;  // :1373
    int len = 0;
    for (int i = 0; i < n; i++) L[len++] = i;
    // splits: the first mspl particles (in element order) whose region exceeds max_w (:1351, :1378-1392)
    int nspl = 0;
    for (int i = 0; i < n && nspl < mspl; i++)
        if (w[p0 + i] > max_w) {
            L[len++] = i | PCDVC_SPLIT;
            nspl++;
        }
    // deletions: candidates below min_w in decreasing swarm index; each deleted slot takes the element's last
    // entry (:1352-1359, :1404-1409; Swarm_RemoveParticleFromCell, SwarmClass.c:1055-1071)
    int ndel = 0;
    for (int i = n - 1; i >= 0 && ndel < mdel; i--) {
        if (!(w[p0 + i] < min_w)) continue;
        int pos = 0;
        while (L[pos] != i) pos++;  // i is in the list: only earlier (larger) candidates were removed
        L[pos] = L[len - 1];
        len--;
        ndel++;
    }
    if ((nspl || ndel) && len < 3) atomicOr(flag, 1);  // :1418-1421
    new_count[e] = len;
    if (nspl) atomicAdd(flag + 1, nspl);
    if (ndel) atomicAdd(flag + 2, ndel);
}

template <int DIM>
__global__ void k_pcdvc_build(int nel, int max_splits, const int *__restrict__ cell_off,
                              const int *__restrict__ new_off, const int *__restrict__ list,
                              const double *__restrict__ x0, const double *__restrict__ x1,
                              const double *__restrict__ x2, const double *__restrict__ c0,
                              const double *__restrict__ c1, const double *__restrict__ c2,
                              const unsigned char *__restrict__ mat, const float *__restrict__ coh,
                              const double *__restrict__ s0, const double *__restrict__ s1, double *__restrict__ o0,
                              double *__restrict__ o1, double *__restrict__ o2, unsigned char *__restrict__ omat,
                              float *__restrict__ ocoh, double *__restrict__ os0, double *__restrict__ os1) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel) return;
    const int p0 = cell_off[e], q0 = new_off[e], len = new_off[e + 1] - q0;
    const int *L = list + p0 + (size_t)e * max_splits;
    for (int k = 0; k < len; k++) {
        const int v = L[k];
        const int src = p0 + (v & (PCDVC_SPLIT - 1)), dst = q0 + k;
        const bool split = (v & PCDVC_SPLIT) != 0;
        // a split copy takes its parent's material point and state and moves to the parent's region centroid
        o0[dst] = split ? c0[src] : x0[src];
        o1[dst] = split ? c1[src] : x1[src];
        if (DIM == 3) o2[dst] = split ? c2[src] : x2[src];
        if (mat) omat[dst] = mat[src];
        if (coh) ocoh[dst] = coh[src];
        if (s0) os0[dst] = s0[src];
        if (s1) os1[dst] = s1[src];
    }
}

extern "C" int ust_k_pcdvc_plan(const ust_grid *gr, const int *cell_off, const double *w, const unsigned char *mat,
                                const ust_pcdvc_opts *o, int *new_count, int *list, int *flag, void *stream) {
    if (!gr || !cell_off || !w || !o || !new_count || !list || !flag) return UST_EINVAL;
    if (o->max_splits < 0) return UST_EINVAL;
    GridDev g = grid_dev(gr);
    if (g.nel <= 0) return UST_OK;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_pcdvc_plan<3>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, cell_off, w, mat, *o, new_count, list, flag);
    else
        hipLaunchKernelGGL(k_pcdvc_plan<2>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, cell_off, w, mat, *o, new_count, list, flag);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_pcdvc_build(int dim, int nel, int max_splits, const int *cell_off, const int *new_off,
                                 const int *list, const double *xi0, const double *xi1, const double *xi2,
                                 const double *c0, const double *c1, const double *c2, const unsigned char *mat,
                                 const float *coh, const double *state0, const double *state1, double *o0,
                                 double *o1, double *o2, unsigned char *omat, float *ocoh, double *ostate0,
                                 double *ostate1, void *stream) {
    if (nel < 0 || max_splits < 0 || !cell_off || !new_off || !list || !xi0 || !xi1 || !c0 || !c1 || !o0 || !o1)
        return UST_EINVAL;
    if (dim == 3 && (!xi2 || !c2 || !o2)) return UST_EINVAL;
    if ((mat && !omat) || (coh && !ocoh) || (state0 && !ostate0) || (state1 && !ostate1)) return UST_EINVAL;
    if (nel == 0) return UST_OK;
    if (dim == 3)
        hipLaunchKernelGGL(k_pcdvc_build<3>, dim3(ust_grid_size(nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), nel, max_splits, cell_off, new_off, list, xi0, xi1, xi2, c0, c1, c2,
                           mat, coh, state0, state1, o0, o1, o2, omat, ocoh, ostate0, ostate1);
    else
        hipLaunchKernelGGL(k_pcdvc_build<2>, dim3(ust_grid_size(nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), nel, max_splits, cell_off, new_off, list, xi0, xi1, xi2, c0, c1, c2,
                           mat, coh, state0, state1, o0, o1, o2, omat, ocoh, ostate0, ostate1);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
