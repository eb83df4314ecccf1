// PIC element assembly for gfx950: cell-sorted particles -> quadrature weights with the
// per-particle rheology fused in, and the batched B^T D B element contraction on FP64 MFMA.
//
// Reference loops restated here:
//   particle loop of _ConstitutiveMatrixCartesian_AssembleElement  ConstitutiveMatrixCartesian.c:277-410
//   rheology chain                                                   RheologyMaterial.c:240-270,
//       MaterialViscosity.c:119-131, Arrhenius.c:156-200, ViscosityLimiter.c:106-125,
//       YieldRheology.c:198-243, VonMises.c:182-269, DruckerPrager.c:236-345, TensorMath.c:430-460
//   G_e: GradientStiffnessMatrixTerm.c:220-249 ; f_e: BuoyancyForceTerm.c:370-448 ;
//   approxS_e: UzawaPreconditionerTerm.c:195-212
#include <math.h>
#include <string.h>

#include "ust_common.h"

// ------------------------------------------------------------------------------------------
// per-particle rheology (isotropic D carried as eta)
// ------------------------------------------------------------------------------------------
// per-particle yield state (strain weakening, tensile flag); pfw == nullptr: none
struct YieldDev {
    double *pfw;
    double *pfw_inc;
    unsigned char *tensile;
    double softening_strain, healing_rate;
};

template <int DIM, bool NEED_T, bool NEED_VEL, bool SW>
__device__ double particle_viscosity(const GridDev &g, const ust_material &M, const int *nodes, const double *xi,
                                     const double *T, const double *vel, const double *pres, int e, float coh,
                                     int prev, const YieldDev &ys, int pi, int *yflag) {
    constexpr int NV = 1 << DIM;
    double eta = 0.0;
    for (int r = 0; r < M.nrh; r++) {
        const ust_rheology &rh = M.rh[r];
        if (rh.type == UST_RH_MATERIALVISCOSITY) {
            eta = rh.p[0];
        } else if (NEED_T && rh.type == UST_RH_ARRHENIUS) {
            double Tp = 0.0, y = 0.0;
#pragma unroll
            for (int a = 0; a < NV; a++) {
                double N = 1.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) N *= 0.5 * (1.0 + sgn((a >> d) & 1) * xi[d]);
                Tp += N * T[nodes[a]];
                y += N * ((a >> 1) & 1);
            }
            if (rh.p[2] > (0.0 + 1e-12)) {
                // y(xi) on a rectilinear element: ymin + (j + local fraction) * hy
                int j = (e / g.nx) % g.ny;
                double yc = g.ymin + ((double)j + y) * g.hy;
                double depth = g.ymax - yc;
                eta = rh.p[0] * exp((rh.p[1] + rh.p[2] * depth) / (Tp + rh.p[3]));
            } else {
                eta = rh.p[0] * exp((rh.p[1]) / (Tp + rh.p[3]));
            }
        } else if (rh.type == UST_RH_VISCOSITYLIMITER) {
            if (eta > rh.p[0]) eta = rh.p[0];
            if (eta < rh.p[1]) eta = rh.p[1];
        } else if (NEED_VEL && (rh.type == UST_RH_VONMISES || rh.type == UST_RH_DRUCKERPRAGER)) {
            const bool dp = rh.type == UST_RH_DRUCKERPRAGER;
            const double minVisc = dp ? rh.p[3] : rh.p[1];
            const int srSoft = (int)(dp ? rh.p[4] : rh.p[2]);
            if (!prev) continue;
            // velocity gradient at xi: L[i][k] = sum_a GNx[k][a] u[a][i]
            double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            const double s[3] = {2.0 / g.hx, 2.0 / g.hy, 2.0 / g.hz};
#pragma unroll
            for (int a = 0; a < NV; a++) {
                double dN[3];
#pragma unroll
                for (int k = 0; k < DIM; k++) {
                    double v = 0.5 * sgn((a >> k) & 1);
#pragma unroll
                    for (int d = 0; d < DIM; d++)
                        if (d != k) v *= 0.5 * (1.0 + sgn((a >> d) & 1) * xi[d]);
                    dN[k] = s[k] * v;
                }
#pragma unroll
                for (int i = 0; i < DIM; i++) {
                    double ui = vel[(size_t)nodes[a] * DIM + i];
#pragma unroll
                    for (int k = 0; k < DIM; k++) L[i][k] += dN[k] * ui;
                }
            }
            double inv;
            if (DIM == 3) {
                double sxx = 2.0 * eta * L[0][0], syy = 2.0 * eta * L[1][1], szz = 2.0 * eta * L[2][2];
                double sxy = eta * 2.0 * (0.5 * (L[0][1] + L[1][0]));
                double sxz = eta * 2.0 * (0.5 * (L[0][2] + L[2][0]));
                double syz = eta * 2.0 * (0.5 * (L[1][2] + L[2][1]));
                inv = 0.5 * (sxx * sxx + syy * syy + szz * szz) + sxy * sxy + sxz * sxz + syz * syz;
            } else {
                double sxx = 2.0 * eta * L[0][0], syy = 2.0 * eta * L[1][1];
                double sxy = eta * 2.0 * (0.5 * (L[0][1] + L[1][0]));
                inv = 0.5 * (sxx * sxx + syy * syy) + sxy * sxy;
            }
            double indicator = sqrt(inv);
            // strain softening of the yield parameters (VonMises.c:194-199, DruckerPrager.c:283-291): the
            // weakening state is clamped to >= 0 where it is read (StrainWeakening.c:491-492)
            const bool sw = SW && rh.p[dp ? 7 : 4] != 0.0;
            double pfw = 0.0, ratio = 0.0;
            if (sw) {
                pfw = ys.pfw[pi];
                if (pfw < 0.0) {
                    pfw = 0.0;
                    ys.pfw[pi] = 0.0;
                }
                ratio = pfw / ys.softening_strain;
                if (ratio > 1.0) ratio = 1.0;
            }
            double cohesion = (rh.p[0] * (1.0 - ratio) + rh.p[dp ? 5 : 3] * ratio) * (double)coh, criterion;
            if (!dp) {
                criterion = cohesion;
            } else {
                const double friction = rh.p[1] * (1.0 - ratio) + rh.p[6] * ratio;
                double pressure = pres[e], phi = atan(friction), k, c;
                if (DIM == 3) {
                    double sinphi = sin(phi), oneOver = 1.0 / (sqrt(3.0) * (3.0 - sinphi));
                    k = 6.0 * sinphi * oneOver;
                    c = 6.0 * cohesion * cos(phi) * oneOver;
                } else {
                    c = cohesion * cos(phi);
                    k = sin(phi);
                }
                criterion = k * pressure + c;
                if (ys.tensile) ys.tensile[pi] = criterion <= 0.0;
                if (criterion < rh.p[2]) criterion = rh.p[2];
            }
            if (sw) {  // increment from the viscosity before yielding (StrainWeakening.c:448-482)
                const double beta = (criterion == 0.0 || indicator == 0.0) ? 0.0 : 1.0 - criterion / indicator;
                ys.pfw_inc[pi] = beta <= 0.0 ? -ys.healing_rate * pfw
                                             : (criterion / eta) * (beta / (1.0 - beta) - ys.healing_rate * pfw);
            }
            *yflag = criterion < indicator;
            if (criterion < indicator) {
                if (srSoft) {
                    eta = 2.0 * criterion * criterion * eta / (criterion * criterion + indicator * indicator);
                } else {
                    double beta = 1.0 - criterion / indicator, corr = -eta * beta;
                    if ((eta + corr) < minVisc) corr = minVisc - eta;
                    eta += corr;
                }
            }
            if (eta < minVisc) eta = minVisc;
        }
    }
    return eta;
}

// ------------------------------------------------------------------------------------------
// per-particle density of BuoyancyForceTerm_AssembleElement (BuoyancyForceTerm.c:391-425):
// rho0 (1 - alpha (T - Tref)), times (1 + beta (depth - depthref)) when beta != 0, less the background
// density of the HydrostaticTerm (HydrostaticTerm.c:267-330) at the particle's height.  full = 0: rho0.
// ------------------------------------------------------------------------------------------
struct BuoyDev {
    ust_buoyancy b;
    int full;
    double top;  // box maximum along the vertical axis
};

__device__ __forceinline__ double hydrostatic_density(const ust_buoyancy &b, double y) {
    double thick, boundary;
    if (b.hs_v != 0) {
        thick = (b.hs_height - b.hs_material_boundary) * b.hs_width / (b.hs_width + b.hs_v * b.hs_time);
        boundary = b.hs_material_boundary +
                   (b.hs_upper_density / b.hs_lower_density) * (b.hs_height - b.hs_material_boundary - thick);
    } else {
        thick = b.hs_height - b.hs_material_boundary;
        boundary = b.hs_material_boundary;
    }
    const double height = thick + boundary, h = b.hs_height - y;
    const double T = h < 0 ? b.hs_T0
                           : b.hs_T0 + b.hs_linear_coefficient * h +
                                 b.hs_exponential_coefficient1 * (1 - exp(-b.hs_exponential_coefficient2 * h));
    double density, alpha;
    if (y > height) {
        density = alpha = 0;
    } else if (y > boundary) {
        density = b.hs_upper_density;
        alpha = b.hs_upper_alpha;
    } else {
        density = b.hs_lower_density;
        alpha = b.hs_lower_alpha;
    }
    return density * (1 - alpha * T);
}

template <int DIM, bool NEED_T>
__device__ double particle_density(const GridDev &g, const BuoyDev &bd, const ust_material &M, const int *nodes,
                                   const double *xi, const double *T, int e) {
    if (!bd.full) return M.density;
    const int idx[3] = {e % g.nx, (e / g.nx) % g.ny, e / (g.nx * g.ny)};
    const double mn[3] = {g.xmin, g.ymin, g.zmin}, h[3] = {g.hx, g.hy, g.hz};
    const int va = bd.b.vertical_axis;
    const double cv = mn[va] + ((double)idx[va] + 0.5 * (xi[va] + 1.0)) * h[va];
    const double depth = fabs(bd.top - cv);
    double temperature = 0.0;
    if (NEED_T) {
#pragma unroll
        for (int a = 0; a < (1 << DIM); a++) {
            double N = 1.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) N *= 0.5 * (1.0 + sgn((a >> d) & 1) * xi[d]);
            temperature += N * T[nodes[a]];
        }
    }
    double background = 0.0;
    if (bd.b.hydrostatic) background = hydrostatic_density(bd.b, mn[1] + ((double)idx[1] + 0.5 * (xi[1] + 1.0)) * h[1]);
    double density = M.density * (1.0 - M.alpha * (temperature - M.ref_temperature));
    if (M.beta != 0) density *= (1.0 + M.beta * (depth - M.ref_depth));
    return density - background;
}

// ------------------------------------------------------------------------------------------
// One particle's contributions to its element's 43 (3D) weight sums: the Lagrange weights of K
// (w detJ eta L_q), the vertex weights of G (w detJ N_b) and of f (w detJ rho N_b).  Shared by both
// particle-weight kernels; the viscosity and yield flag go to eta_out / yield_out.
// ------------------------------------------------------------------------------------------
template <int DIM, bool NEED_T, bool NEED_VEL, bool SW>
__device__ __forceinline__ void particle_contrib(
    const GridDev &g, int e, int pi, const int *nodes, double detJ, const double *__restrict__ xi0,
    const double *__restrict__ xi1, const double *__restrict__ xi2, const double *__restrict__ w,
    const unsigned char *__restrict__ mat, const float *__restrict__ coh, const ust_material *__restrict__ mats,
    int nmat, const double *__restrict__ T, const double *__restrict__ vel, const double *__restrict__ pres, int prev,
    const YieldDev &ys, const BuoyDev &bd, double *__restrict__ eta_out, unsigned char *__restrict__ yield_out,
    double *acc) {
    constexpr int NQ = DIM == 3 ? 27 : 9;
    constexpr int NV = 1 << DIM;
    double xi[3];
    xi[0] = xi0[pi];
    xi[1] = xi1[pi];
    xi[2] = DIM == 3 ? xi2[pi] : 0.0;
    const int m = mat ? (int)mat[pi] : 0;
    const ust_material &M = mats[m < nmat ? m : 0];
    int yf = -1;
    const double eta = particle_viscosity<DIM, NEED_T, NEED_VEL, SW>(g, M, nodes, xi, T, vel, pres, e,
                                                                 coh ? coh[pi] : 1.0f, prev, ys, pi, &yf);
    if (eta_out) eta_out[pi] = eta;
    if (NEED_VEL && yield_out && yf >= 0) yield_out[pi] = (unsigned char)yf;
    const double wd = w[pi] * detJ;
    const double et = wd * eta;
    const double wr = wd * particle_density<DIM, NEED_T>(g, bd, M, nodes, xi, T, e);
    double Lx[3], Ly[3], Lz[3];
    lagrange3(xi[0], Lx);
    lagrange3(xi[1], Ly);
    if (DIM == 3) lagrange3(xi[2], Lz);
#pragma unroll
    for (int qz = 0; qz < (DIM == 3 ? 3 : 1); qz++)
#pragma unroll
        for (int qy = 0; qy < 3; qy++) {
            const double eyz = et * Ly[qy] * (DIM == 3 ? Lz[qz] : 1.0);
#pragma unroll
            for (int qx = 0; qx < 3; qx++) acc[qx + 3 * qy + 9 * qz] += eyz * Lx[qx];
        }
    const double Nx[2] = {0.5 * (1.0 - xi[0]), 0.5 * (1.0 + xi[0])};
    const double Ny[2] = {0.5 * (1.0 - xi[1]), 0.5 * (1.0 + xi[1])};
    const double Nz[2] = {0.5 * (1.0 - xi[2]), 0.5 * (1.0 + xi[2])};
#pragma unroll
    for (int b = 0; b < NV; b++) {
        double N = Nx[b & 1] * Ny[(b >> 1) & 1];
        if (DIM == 3) N *= Nz[(b >> 2) & 1];
        acc[NQ + b] += wd * N;
        acc[NQ + NV + b] += wr * N;
    }
}

__device__ __forceinline__ void element_nodes_of(const GridDev &g, int ie, int je, int ke, int dim, int *nodes) {
    const int n0 = ie + g.nnx * (je + g.nny * ke);
    for (int a = 0; a < (1 << dim); a++)
        nodes[a] = n0 + (a & 1) + g.nnx * ((a >> 1) & 1) + (dim == 3 ? g.nnx * g.nny * ((a >> 2) & 1) : 0);
}

// ------------------------------------------------------------------------------------------
// particles -> weights.  A wavefront covers 64/S consecutive elements (S chosen near the mean particle
// count, so few lanes idle); lane k of a segment takes particles k, k+S, ... of its element (coalesced:
// particles are cell-sorted) and accumulates its 43 (3D) partial sums in registers.  The partials are
// then reduced through LDS in chunks of PW_CH columns: every lane writes its chunk, and lane t sums
// column t of one segment; the odd row length keeps the 64-bit row writes of a wavefront free of bank
// conflicts, and each wave reduces in its own LDS rows, so waves never wait for each other.  Workgroups
// are persistent (grid-stride over element groups) and copy the material table
// (at most UST_MAX_MATERIALS entries) into LDS once, so the per-particle rheology chain reads LDS instead
// of making dependent global loads.  NEED_T / NEED_VEL / SW compile the temperature and strain-rate lookups
// (Arrhenius, yield) and the strain-weakening state in only when a material uses them.
// ------------------------------------------------------------------------------------------
#define PW_WAVES 2
#define PW_CH 11
// a wavefront's own LDS rows: its LDS operations complete in order, so a wave-level barrier (no compiler
// reordering across it) is all the reduction needs between a wave's writes and its reads
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <int DIM, int S, bool NEED_T, bool NEED_VEL, bool SW>
__global__ __launch_bounds__(64 * PW_WAVES, (DIM == 2 ? 2 : 1)) void k_particle_weights(
    GridDev g, const int *__restrict__ cell_off, const double *__restrict__ xi0, const double *__restrict__ xi1,
    const double *__restrict__ xi2, const double *__restrict__ w, const unsigned char *__restrict__ mat,
    const float *__restrict__ coh, const ust_material *__restrict__ mats, int nmat, const double *__restrict__ T,
    const double *__restrict__ vel, const double *__restrict__ pres, int prev, YieldDev ys, BuoyDev bd,
    double *__restrict__ eta_out, unsigned char *__restrict__ yield_out, double *__restrict__ WK,
    double *__restrict__ WG, double *__restrict__ WF) {
    constexpr int NQ = DIM == 3 ? 27 : 9;
    constexpr int NV = 1 << DIM;
    constexpr int NA = NQ + 2 * NV;  // accumulators per lane
    constexpr int LD = PW_CH;        // odd row length (in doubles)
    constexpr int nseg = 64 / S;
    __shared__ double red[PW_WAVES][64 * LD];
    __shared__ size_t sbase[PW_WAVES][nseg];
    __shared__ ust_material smat[UST_MAX_MATERIALS];
    {
        const int nw = nmat * (int)(sizeof(ust_material) / sizeof(int));
        const int *src = (const int *)mats;
        int *dst = (int *)smat;
        for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int seg = lane / S, k = lane % S;
    const int ngroups = (g.nel + PW_WAVES * nseg - 1) / (PW_WAVES * nseg);
    const double detJ = (DIM == 3) ? 0.125 * g.hx * g.hy * g.hz : 0.25 * g.hx * g.hy;
    double *r = red[wv];
    // the next group's particle range is loaded one group ahead, while this group is reduced
    int grp = blockIdx.x;
    int p0 = 0, p1 = 0;
    {
        const int e = (grp * PW_WAVES + wv) * nseg + seg;
        if (grp < ngroups && seg < nseg && e < g.nel) {
            p0 = cell_off[e];
            p1 = cell_off[e + 1];
        }
    }
    for (; grp < ngroups; grp += gridDim.x) {
        const int ebase = (grp * PW_WAVES + wv) * nseg;
        const int e = ebase + seg;
        const int pbeg = p0, pend = p1;
        {
            const int en = ebase + gridDim.x * PW_WAVES * nseg + seg;
            if (grp + (int)gridDim.x < ngroups && seg < nseg && en < g.nel) {
                p0 = cell_off[en];
                p1 = cell_off[en + 1];
            }
        }
        double acc[NA];
#pragma unroll
        for (int t = 0; t < NA; t++) acc[t] = 0.0;
        // slot of K weight 0 of each segment's element (weight q sits q * UST_TILE further on in every layout),
        // left in LDS for the reduction threads instead of recomputing the layout index per store
        if (seg < nseg && e < g.nel) {
            if (k == 0) sbase[wv][seg] = widx(g, e, 0);
            const int ie = e % g.nx, je = (e / g.nx) % g.ny, ke = e / (g.nx * g.ny);
            int nodes[NV];
            element_nodes_of(g, ie, je, ke, DIM, nodes);
            for (int pi = pbeg + k; pi < pend; pi += S)
                particle_contrib<DIM, NEED_T, NEED_VEL, SW>(g, e, pi, nodes, detJ, xi0, xi1, xi2, w, mat, coh, smat,
                                                            nmat, T, vel, pres, prev, ys, bd, eta_out, yield_out,
                                                            acc);
        }
#pragma unroll
        for (int c0 = 0; c0 < NA; c0 += PW_CH) {
#pragma unroll
            for (int t = 0; t < PW_CH; t++)
                if (c0 + t < NA) r[lane * LD + t] = acc[c0 + t];
            wave_sync();
            const int nc = NA - c0 < PW_CH ? NA - c0 : PW_CH;
            for (int t = lane; t < nseg * nc; t += 64) {
                const int sg = t / nc, col = c0 + t % nc, el = ebase + sg;
                if (el < g.nel) {
                    // S independent LDS reads into four partial sums (S is a compile-time constant)
                    double ps[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int row = 0; row < S; row++) ps[row & 3] += r[(sg * S + row) * LD + (col - c0)];
                    const double sum = (ps[0] + ps[1]) + (ps[2] + ps[3]);
                    if (col < NQ) WK[sbase[wv][sg] + (size_t)col * UST_TILE] = sum;
                    else if (col < NQ + NV) WG[(size_t)(col - NQ) * g.nel + el] = sum;
                    else WF[(size_t)(col - NQ - NV) * g.nel + el] = sum;
                }
            }
            wave_sync();
        }
    }
}

// ------------------------------------------------------------------------------------------
// B(xi_q)^T D0 B(xi_q) entry (row (a,i), col (b,j)) on the Lagrange point q, from the global
// shape-function derivatives GNx_k(a) at q:  delta_ij sum_k GNx_k(a) GNx_k(b) + GNx_j(a) GNx_i(b)
// ------------------------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ double gnx_at(int a, int k, int q, const double *s) {
    // s_k * dN_a/dxi_k at Lagrange point q:  sgn(a_k) prod_{d != k} cfac(a_d, q_d) / 2^DIM
    int qd[3] = {q % 3, (q / 3) % 3, q / 9};
    int v = sgn((a >> k) & 1);
#pragma unroll
    for (int d = 0; d < DIM; d++)
        if (d != k) v *= cfac((a >> d) & 1, qd[d]);
    return s[k] * (double)v / (double)(1 << DIM);
}

typedef double ust_double4 __attribute__((ext_vector_type(4)));

// K_e = W (count x NQ) * M (NQ x nd^2) on v_mfma_f64_16x16x4f64.  One wavefront per 16 elements;
// A fragment: lane l holds A[l%16][l/16], B fragment: B[l/16][l%16], D: acc[r] = D[4r + l/16][l%16]
// (output layout measured on gfx950 against the oracle's element matrices).  The B operand M[q][(a,i),(b,j)]
// is built from the shape-function derivatives at the Lagrange points, GN[a][k][q] (2^DIM x DIM x NQ values),
// which the workgroup tabulates in LDS once: an entry of M is then at most four products of table values.
template <int DIM>
__global__ __launch_bounds__(256) void k_element_matrices(GridDev g, const double *__restrict__ WK, int first,
                                                          int count, double *__restrict__ Ke) {
    constexpr int NQ = DIM == 3 ? 27 : 9;
    constexpr int NV = 1 << DIM;
    constexpr int ND = NV * DIM;
    constexpr int NCOL = ND * ND;
    constexpr int KSTEPS = (NQ + 3) / 4;
    constexpr int NG = NV * DIM * NQ;
    __shared__ double gn[NG];  // gn[(a*DIM + k)*NQ + q]
    const double s[3] = {2.0 / g.hx, 2.0 / g.hy, 2.0 / g.hz};
    for (int t = threadIdx.x; t < NG; t += blockDim.x) {
        const int q = t % NQ, ak = t / NQ;
        gn[t] = gnx_at<DIM>(ak / DIM, ak % DIM, q, s);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int e0 = tile * 16;
    if (e0 >= count) return;
    const int li = lane & 15, lk = lane >> 4;
    double afrag[KSTEPS];
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ks++) {
        const int q = ks * 4 + lk, el = e0 + li;
        afrag[ks] = (q < NQ && el < count) ? WK[widx(g, first + el, q)] : 0.0;
    }
    for (int ct = 0; ct < NCOL / 16; ct++) {
        ust_double4 acc = {0.0, 0.0, 0.0, 0.0};
        const int col = ct * 16 + li;
        const int row = col / ND, cin = col % ND;
        const int a = row / DIM, i = row % DIM, b = cin / DIM, j = cin % DIM;
        const double *ga = gn + a * DIM * NQ, *gb = gn + b * DIM * NQ;
#pragma unroll
        for (int ks = 0; ks < KSTEPS; ks++) {
            const int q = ks * 4 + lk;
            double m = 0.0;
            if (q < NQ) {
                if (i == j) {
#pragma unroll
                    for (int k = 0; k < DIM; k++) m += ga[k * NQ + q] * gb[k * NQ + q];
                }
                m += ga[j * NQ + q] * gb[i * NQ + q];
            }
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[ks], m, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int el = e0 + 4 * r + lk;
            if (el < count) Ke[(size_t)el * NCOL + col] = acc[r];
        }
    }
}

// G_e[a,i] = -(sgn(a_i)/h_i) (WG[b_i=0,b_perp=a_perp] + WG[b_i=1,b_perp=a_perp])
template <int DIM>
__device__ __forceinline__ void element_G(const GridDev &g, const double *WG, int e, double *Ge) {
    constexpr int NV = 1 << DIM;
    double wv[NV];
#pragma unroll
    for (int b = 0; b < NV; b++) wv[b] = WG[(size_t)b * g.nel + e];
    const double h[3] = {g.hx, g.hy, g.hz};
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            const int b0 = a & ~(1 << i), b1 = a | (1 << i);
            Ge[a * DIM + i] = -(double)sgn((a >> i) & 1) / h[i] * (wv[b0] + wv[b1]);
        }
}

template <int DIM>
__global__ void k_element_G(GridDev g, const double *__restrict__ WG, int first, int count, double *__restrict__ Ge) {
    constexpr int ND = (1 << DIM) * DIM;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    double ge[ND];
    element_G<DIM>(g, WG, first + t, ge);
#pragma unroll
    for (int v = 0; v < ND; v++) Ge[(size_t)t * ND + v] = ge[v];
}

struct FDir {
    double d[3];
};

template <int DIM>
__global__ void k_element_f(GridDev g, const double *__restrict__ WF, double gravity, FDir dir, int first, int count,
                            double *__restrict__ fe) {
    constexpr int NV = 1 << DIM, ND = NV * DIM;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int i = 0; i < DIM; i++)
            fe[(size_t)t * ND + a * DIM + i] = dir.d[i] != 0.0 ? gravity * dir.d[i] * WF[(size_t)a * g.nel + first + t]
                                                                : 0.0;
}

// BuoyancyForceTerm "adjust": WF *= 2^dim detJ / sum_b WG[b] (= 2^dim / the element's sum of weights)
template <int DIM>
__global__ void k_adjust_force(GridDev g, const double *__restrict__ WG, double *__restrict__ WF) {
    constexpr int NV = 1 << DIM;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const double detJ = (DIM == 3) ? 0.125 * g.hx * g.hy * g.hz : 0.25 * g.hx * g.hy;
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NV; b++) s += WG[(size_t)b * g.nel + e];
    const double f = s > 0.0 ? (double)NV * detJ / s : 1.0;
#pragma unroll
    for (int b = 0; b < NV; b++) WF[(size_t)b * g.nel + e] *= f;
}

// diag of K_e: sum_q W_q (sum_k GNx_k(a)^2 + GNx_i(a)^2)
template <int DIM>
__device__ __forceinline__ void element_K_diag(const GridDev &g, const double *WK, int e, double *dg) {
    constexpr int NQ = DIM == 3 ? 27 : 9, NV = 1 << DIM;
    const double s[3] = {2.0 / g.hx, 2.0 / g.hy, 2.0 / g.hz};
#pragma unroll
    for (int v = 0; v < NV * DIM; v++) dg[v] = 0.0;
    for (int q = 0; q < NQ; q++) {
        const double wq = WK[widx(g, e, q)];
#pragma unroll
        for (int a = 0; a < NV; a++) {
            double gk[3], s2 = 0.0;
#pragma unroll
            for (int k = 0; k < DIM; k++) {
                gk[k] = gnx_at<DIM>(a, k, q, s);
                s2 += gk[k] * gk[k];
            }
#pragma unroll
            for (int i = 0; i < DIM; i++) dg[a * DIM + i] += wq * (s2 + gk[i] * gk[i]);
        }
    }
}

template <int DIM>
__global__ void k_uzawa_diag(GridDev g, const double *__restrict__ WK, const double *__restrict__ WG,
                             double *__restrict__ approxS) {
    constexpr int ND = (1 << DIM) * DIM;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    double dg[ND], ge[ND];
    element_K_diag<DIM>(g, WK, e, dg);
    element_G<DIM>(g, WG, e, ge);
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < ND; v++) s += ge[v] * (ge[v] / dg[v]);
    approxS[e] = s;
}

// ------------------------------------------------------------------------------------------
extern "C" int ust_k_particle_weights(const ust_grid *gr, const int *cell_off, int ppc_hint, const double *xi0,
                                      const double *xi1, const double *xi2, const double *w,
                                      const unsigned char *mat, const float *coh_scale,
                                      const ust_material *materials_dev, int nmat, const double *temperature,
                                      const double *velocity, const double *pressure, int prev_solution,
                                      const ust_yield_state *yield_state, const ust_buoyancy *buoyancy,
                                      double *eta_out, unsigned char *yield_out, double *WK, double *WG, double *WF,
                                      void *stream) {
    if (!gr || !cell_off || !xi0 || !xi1 || !w || !materials_dev || !WK || !WG || !WF) return UST_EINVAL;
    if (gr->dim == 3 && !xi2) return UST_EINVAL;
    if (nmat < 1 || nmat > UST_MAX_MATERIALS) return UST_EINVAL;
    GridDev g = grid_dev(gr);
    if (g.nel <= 0) return UST_OK;
    BuoyDev bd;
    memset(&bd, 0, sizeof(bd));
    if (buoyancy) {
        if (buoyancy->vertical_axis < 0 || buoyancy->vertical_axis >= gr->dim) return UST_EINVAL;
        bd.b = *buoyancy;
        bd.full = 1;
        bd.top = gr->xmax[buoyancy->vertical_axis];
    }
    YieldDev ys = {nullptr, nullptr, nullptr, 1.0, 0.0};
    if (yield_state) {
        if (yield_state->pfw && (!yield_state->pfw_inc || !(yield_state->softening_strain > 0.0))) return UST_EINVAL;
        ys.pfw = yield_state->pfw;
        ys.pfw_inc = yield_state->pfw_inc;
        ys.tensile = yield_state->tensile;
        ys.softening_strain = yield_state->softening_strain;
        ys.healing_rate = yield_state->healing_rate;
    }
    // temperature / velocity lookups only when a rheology or the density reads them (the caller passes NULL
    // otherwise)
    const bool needT = temperature != nullptr, needV = velocity != nullptr && pressure != nullptr;
    hipStream_t st = ust_stream(stream);
    // segment width: the smallest candidate >= the caller's particles-per-cell hint (the mean count; a cell
    // with more particles than S loops), so few lanes idle; segments never straddle a wavefront
    int S = 64;
    const int cands[6] = {8, 12, 16, 21, 32, 64};
    for (int c = 0; c < 6; c++)
        if (cands[c] >= ppc_hint) {
            S = cands[c];
            break;
        }
    const int ngroups = (g.nel + PW_WAVES * (64 / S) - 1) / (PW_WAVES * (64 / S));
    // persistent workgroups: as many as fit on the device at once (at least one per compute unit)
    static int n_cu = 0;
    if (n_cu == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n_cu = v;
        else
            n_cu = 256;
    }
#define LAUNCH_PW(D, SS, NT, NVL, SWK)                                                                         \
    do {                                                                                                       \
        int per_cu = 0;                                                                                        \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_particle_weights<D, SS, NT, NVL, SWK>,     \
                                                         64 * PW_WAVES, 0) != hipSuccess || per_cu < 1)        \
            per_cu = 1;                                                                                        \
        const int nb = ngroups < n_cu * per_cu ? ngroups : n_cu * per_cu;                                      \
        hipLaunchKernelGGL((k_particle_weights<D, SS, NT, NVL, SWK>), dim3(nb), dim3(64 * PW_WAVES), 0, st, g, \
                           cell_off, xi0, xi1, xi2, w, mat, coh_scale, materials_dev, nmat, temperature,       \
                           velocity, pressure, prev_solution, ys, bd, eta_out, yield_out, WK, WG, WF);         \
    } while (0)
#define LAUNCH_S(D, NT, NVL, SWK)                                  \
    do {                                                           \
        switch (S) {                                               \
            case 8: LAUNCH_PW(D, 8, NT, NVL, SWK); break;          \
            case 12: LAUNCH_PW(D, 12, NT, NVL, SWK); break;        \
            case 16: LAUNCH_PW(D, 16, NT, NVL, SWK); break;        \
            case 21: LAUNCH_PW(D, 21, NT, NVL, SWK); break;        \
            case 32: LAUNCH_PW(D, 32, NT, NVL, SWK); break;        \
            default: LAUNCH_PW(D, 64, NT, NVL, SWK); break;        \
        }                                                          \
    } while (0)
    // strain weakening is compiled in only when the yield state carries it (it costs registers)
    const bool sw = needV && ys.pfw != nullptr;
#define LAUNCH_F(D)                                                  \
    do {                                                             \
        if (needT && needV && sw) LAUNCH_S(D, true, true, true);     \
        else if (needT && needV) LAUNCH_S(D, true, true, false);     \
        else if (needT) LAUNCH_S(D, true, false, false);             \
        else if (needV && sw) LAUNCH_S(D, false, true, true);        \
        else if (needV) LAUNCH_S(D, false, true, false);             \
        else LAUNCH_S(D, false, false, false);                       \
    } while (0)
    if (g.dim == 3) LAUNCH_F(3);
    else LAUNCH_F(2);
#undef LAUNCH_F
#undef LAUNCH_S
#undef LAUNCH_PW
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_element_matrices(const ust_grid *gr, const double *WK, int first, int count, double *Ke,
                                      void *stream) {
    GridDev g = grid_dev(gr);
    if (first < 0 || count < 0 || first + count > g.nel || !WK || !Ke) return UST_EINVAL;
    if (count == 0) return UST_OK;
    const int tiles = (count + 15) / 16;
    const unsigned grid = ust_grid_size(tiles, UST_BLOCK / 64);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_element_matrices<3>, dim3(grid), dim3(UST_BLOCK), 0, ust_stream(stream), g, WK, first,
                           count, Ke);
    else
        hipLaunchKernelGGL(k_element_matrices<2>, dim3(grid), dim3(UST_BLOCK), 0, ust_stream(stream), g, WK, first,
                           count, Ke);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_element_G(const ust_grid *gr, const double *WG, int first, int count, double *Ge,
                               void *stream) {
    GridDev g = grid_dev(gr);
    if (first < 0 || count < 0 || first + count > g.nel || !WG || !Ge) return UST_EINVAL;
    if (count == 0) return UST_OK;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_element_G<3>, dim3(ust_grid_size(count, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WG, first, count, Ge);
    else
        hipLaunchKernelGGL(k_element_G<2>, dim3(ust_grid_size(count, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WG, first, count, Ge);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_element_f_dir(const ust_grid *gr, const double *WF, double gravity, const double *dir,
                                   int first, int count, double *fe, void *stream) {
    GridDev g = grid_dev(gr);
    if (first < 0 || count < 0 || first + count > g.nel || !WF || !fe || !dir) return UST_EINVAL;
    if (count == 0) return UST_OK;
    FDir d = {{dir[0], dir[1], g.dim == 3 ? dir[2] : 0.0}};
    if (g.dim == 3)
        hipLaunchKernelGGL(k_element_f<3>, dim3(ust_grid_size(count, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WF, gravity, d, first, count, fe);
    else
        hipLaunchKernelGGL(k_element_f<2>, dim3(ust_grid_size(count, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WF, gravity, d, first, count, fe);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_element_f(const ust_grid *gr, const double *WF, double gravity, int first, int count,
                               double *fe, void *stream) {
    const double down[3] = {0.0, -1.0, 0.0};
    return ust_k_element_f_dir(gr, WF, gravity, down, first, count, fe, stream);
}

extern "C" int ust_k_adjust_force_weights(const ust_grid *gr, const double *WG, double *WF, void *stream) {
    GridDev g = grid_dev(gr);
    if (!WG || !WF) return UST_EINVAL;
    if (g.nel == 0) return UST_OK;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_adjust_force<3>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WG, WF);
    else
        hipLaunchKernelGGL(k_adjust_force<2>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WG, WF);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" int ust_k_uzawa_diag(const ust_grid *gr, const double *WK, const double *WG, double *approxS,
                                void *stream) {
    GridDev g = grid_dev(gr);
    if (!WK || !WG || !approxS) return UST_EINVAL;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_uzawa_diag<3>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WK, WG, approxS);
    else
        hipLaunchKernelGGL(k_uzawa_diag<2>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WK, WG, approxS);
    UST_LAUNCH_CHECK();
    return UST_OK;
}
