// Matrix-free Stokes operators for gfx950 on the quadrature-weight representation:
// K u (the MatMult of the A11 block inside BSSCR and its multigrid smoothers,
// auglag-driver-DGTGD.c:191-403, PETScMGSolver.c:641-716), G p and G^T u
// (MatSchurComplement's A01/A10, StokesBlockKSPInterface.c:242), BC lifting
// (StiffnessMatrix.c:900-951), Galerkin coarse operators (PETScMGSolver.c:518-527) and the
// SROpGenerator transfers (SROpGenerator.c:1000-1041).
#include <math.h>

#include "ust_common.h"

// ------------------------------------------------------------------------------------------
// Element action y_e = K_e u_e with K_e = sum_q W_q B_q^T D0 B_q.  Gradients are taken at the
// 3^DIM Lagrange points, tau = W_q (grad u + grad u^T), y_a,i = sum_q sum_k GNx_k(a) tau_ik.
// All coefficients are compile-time constants after unrolling (values 0, +-1, +-2 / 2^DIM).
// ------------------------------------------------------------------------------------------
template <int DIM, typename T>
__device__ __forceinline__ void element_apply(const T *wq, const T s[3], const T u[1 << DIM][DIM],
                                              T y[1 << DIM][DIM]) {
    constexpr int NV = 1 << DIM;
    constexpr int NQ = DIM == 3 ? 27 : 9;
    const T inv = (T)1.0 / (T)(1 << DIM);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int qd[3] = {q % 3, (q / 3) % 3, q / 9};
        T gr[DIM][DIM];  // gr[i][k] = d u_i / d x_k
#pragma unroll
        for (int i = 0; i < DIM; i++)
#pragma unroll
            for (int k = 0; k < DIM; k++) gr[i][k] = (T)0;
#pragma unroll
        for (int a = 0; a < NV; a++) {
#pragma unroll
            for (int k = 0; k < DIM; k++) {
                int c = sgn((a >> k) & 1);
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (d != k) c *= cfac((a >> d) & 1, qd[d]);
                if (c != 0) {
#pragma unroll
                    for (int i = 0; i < DIM; i++) gr[i][k] += (T)c * u[a][i];
                }
            }
        }
        T tau[DIM][DIM];
        const T wv = wq[q] * inv;
#pragma unroll
        for (int i = 0; i < DIM; i++)
#pragma unroll
            for (int k = 0; k < DIM; k++) tau[i][k] = wv * (s[k] * gr[i][k] + s[i] * gr[k][i]);
#pragma unroll
        for (int a = 0; a < NV; a++) {
#pragma unroll
            for (int k = 0; k < DIM; k++) {
                int c = sgn((a >> k) & 1);
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (d != k) c *= cfac((a >> d) & 1, qd[d]);
                if (c != 0) {
                    const T cs = (T)c * s[k] * inv;
#pragma unroll
                    for (int i = 0; i < DIM; i++) y[a][i] += cs * tau[i][k];
                }
            }
        }
    }
}

// INMODE 0: input on free dofs (normal apply);  1: input on BC dofs only (lifting).
// SIGN: +1 y += K u ; -1 y -= K u.  Output on free dofs only, accumulated with atomics.
template <int DIM, typename T, int INMODE, int SIGN>
__global__ __launch_bounds__(256) void k_apply_K(GridDev g, const T *__restrict__ W,
                                                 const unsigned char *__restrict__ mask, const T *__restrict__ u,
                                                 T *__restrict__ y) {
    constexpr int NV = 1 << DIM;
    constexpr int NQ = DIM == 3 ? 27 : 9;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int ie = e % g.nx, je = (e / g.nx) % g.ny, ke = e / (g.nx * g.ny);
    const int n0 = ie + g.nnx * (je + g.nny * ke);
    int nodes[NV];
#pragma unroll
    for (int a = 0; a < NV; a++)
        nodes[a] = n0 + (a & 1) + g.nnx * ((a >> 1) & 1) + (DIM == 3 ? g.nnx * g.nny * ((a >> 2) & 1) : 0);
    T ue[NV][DIM], ye[NV][DIM], wq[NQ];
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            const size_t dof = (size_t)nodes[a] * DIM + i;
            const bool bc = mask[dof] != 0;
            ue[a][i] = ((INMODE == 0) ? !bc : bc) ? u[dof] : (T)0;
            ye[a][i] = (T)0;
        }
#pragma unroll
    for (int q = 0; q < NQ; q++) wq[q] = W[widx(g, e, q)];
    const T s[3] = {(T)(2.0 / g.hx), (T)(2.0 / g.hy), (T)(2.0 / g.hz)};
    element_apply<DIM, T>(wq, s, ue, ye);
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            const size_t dof = (size_t)nodes[a] * DIM + i;
            if (!mask[dof]) unsafeAtomicAdd(&y[dof], SIGN > 0 ? ye[a][i] : -ye[a][i]);
        }
}

template <typename T>
__global__ void k_fill(T *x, T v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = v;
}

template <int DIM>
__device__ __forceinline__ void elem_nodes(const GridDev &g, int e, int *nodes) {
    const int ie = e % g.nx, je = (e / g.nx) % g.ny, ke = e / (g.nx * g.ny);
    const int n0 = ie + g.nnx * (je + g.nny * ke);
#pragma unroll
    for (int a = 0; a < (1 << DIM); a++)
        nodes[a] = n0 + (a & 1) + g.nnx * ((a >> 1) & 1) + (DIM == 3 ? g.nnx * g.nny * ((a >> 2) & 1) : 0);
}

// G_e[a,i] = -(sgn(a_i)/h_i)(WG[a with bit i = 0] + WG[a with bit i = 1])
template <int DIM>
__device__ __forceinline__ double g_entry(const GridDev &g, const double *wv, int a, int i) {
    const double h = i == 0 ? g.hx : (i == 1 ? g.hy : g.hz);
    return -(double)sgn((a >> i) & 1) / h * (wv[a & ~(1 << i)] + wv[a | (1 << i)]);
}

// one thread per node: y[n, i] = sum over the <= 2^DIM elements around n of G_e[a, i] p_e, a being the
// node's vertex index in e.  Every dof is written once (0 on Dirichlet dofs) in a fixed order.
template <int DIM, bool ABS>
__global__ __launch_bounds__(256) void k_apply_G(GridDev g, const double *__restrict__ WG,
                                                 const unsigned char *__restrict__ mask, const double *__restrict__ p,
                                                 double *__restrict__ y) {
    constexpr int NV = 1 << DIM;
    const size_t nnodes = (size_t)g.nnx * g.nny * g.nnz;
    const size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    const int gx = (int)(n % g.nnx), gy = (int)((n / g.nnx) % g.nny), gz = (int)(n / ((size_t)g.nnx * g.nny));
    double acc[DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++) acc[i] = 0.0;
#pragma unroll
    for (int a = 0; a < NV; a++) {
        const int ex = gx - (a & 1), ey = gy - ((a >> 1) & 1), ez = DIM == 3 ? gz - (a >> 2) : 0;
        if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
        const size_t e = (size_t)ex + (size_t)g.nx * ((size_t)ey + (size_t)g.ny * ez);
        const double pe = ABS ? 1.0 : p[e], wa = WG[(size_t)a * g.nel + e];
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            const double h = i == 0 ? g.hx : (i == 1 ? g.hy : g.hz);
            const double gai = -(double)sgn((a >> i) & 1) / h * (wa + WG[(size_t)(a ^ (1 << i)) * g.nel + e]);
            acc[i] += ABS ? fabs(gai) : gai * pe;
        }
    }
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        const size_t dof = n * DIM + i;
        y[dof] = mask[dof] ? 0.0 : acc[i];
    }
}

// MODE 0: q_e = sum_free G_e u ;  MODE 1: q_e -= sum_bc G_e u  (continuity-RHS lifting)
template <int DIM, int MODE>
__global__ __launch_bounds__(256) void k_apply_Gt(GridDev g, const double *__restrict__ WG,
                                                  const unsigned char *__restrict__ mask,
                                                  const double *__restrict__ u, double *__restrict__ q) {
    constexpr int NV = 1 << DIM;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    int nodes[NV];
    elem_nodes<DIM>(g, e, nodes);
    double wv[NV];
#pragma unroll
    for (int b = 0; b < NV; b++) wv[b] = WG[(size_t)b * g.nel + e];
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            const size_t dof = (size_t)nodes[a] * DIM + i;
            const bool bc = mask[dof] != 0;
            if (MODE == 2) {
                if (!bc) s += fabs(g_entry<DIM>(g, wv, a, i));
            } else if (MODE == 3) {
                if (!bc) {
                    const double gi = g_entry<DIM>(g, wv, a, i);
                    s += gi * gi * u[dof];
                }
            } else if (MODE == 0 ? !bc : bc) {
                s += g_entry<DIM>(g, wv, a, i) * u[dof];
            }
        }
    if (MODE == 1) q[e] -= s;
    else q[e] = s;
}

struct Dir3 {
    double d[3];
};

// one thread per node: f[n, c] += g dir[c] sum over the elements around n of WF[a][e]  (free dofs only; dir is
// -1 on the vertical axis, or the gravity direction gHat)
template <int DIM>
__global__ void k_force(GridDev g, const double *__restrict__ WF, double gravity, Dir3 dir,
                        const unsigned char *__restrict__ mask, double *__restrict__ f) {
    constexpr int NV = 1 << DIM;
    const size_t nnodes = (size_t)g.nnx * g.nny * g.nnz;
    const size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (n >= nnodes) return;
    const int gx = (int)(n % g.nnx), gy = (int)((n / g.nnx) % g.nny), gz = (int)(n / ((size_t)g.nnx * g.nny));
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < NV; a++) {
        const int ex = gx - (a & 1), ey = gy - ((a >> 1) & 1), ez = DIM == 3 ? gz - (a >> 2) : 0;
        if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
        acc += WF[(size_t)a * g.nel + (size_t)ex + (size_t)g.nx * ((size_t)ey + (size_t)g.ny * ez)];
    }
#pragma unroll
    for (int c = 0; c < DIM; c++) {
        const size_t dof = n * DIM + c;
        if (dir.d[c] != 0.0 && !mask[dof]) f[dof] += gravity * dir.d[c] * acc;
    }
}

template <int DIM>
__global__ void k_diag_acc(GridDev g, const double *__restrict__ WK, const unsigned char *__restrict__ mask,
                           double *__restrict__ dg) {
    constexpr int NV = 1 << DIM, NQ = DIM == 3 ? 27 : 9;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    int nodes[NV];
    elem_nodes<DIM>(g, e, nodes);
    const double s[3] = {2.0 / g.hx, 2.0 / g.hy, 2.0 / g.hz};
    double acc[NV][DIM];
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int i = 0; i < DIM; i++) acc[a][i] = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int qd[3] = {q % 3, (q / 3) % 3, q / 9};
        const double wq = WK[widx(g, e, q)] / (double)(1 << (2 * DIM));
#pragma unroll
        for (int a = 0; a < NV; a++) {
            double gk[3], s2 = 0.0;
#pragma unroll
            for (int k = 0; k < DIM; k++) {
                int c = sgn((a >> k) & 1);
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (d != k) c *= cfac((a >> d) & 1, qd[d]);
                gk[k] = s[k] * c;
                s2 += gk[k] * gk[k];
            }
#pragma unroll
            for (int i = 0; i < DIM; i++) acc[a][i] += wq * (s2 + gk[i] * gk[i]);
        }
    }
#pragma unroll
    for (int a = 0; a < NV; a++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            const size_t dof = (size_t)nodes[a] * DIM + i;
            if (!mask[dof]) unsafeAtomicAdd(&dg[dof], acc[a][i]);
        }
}

__global__ void k_diag_invert(const unsigned char *__restrict__ mask, double *__restrict__ d64,
                              float *__restrict__ d32, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = mask[i] ? 1.0 : 1.0 / d64[i];
    d64[i] = v;
    if (d32) d32[i] = (float)v;
}

// ------------------------------------------------------------------------------------------
// Galerkin coarsening of the weights: W^E_{q'} = sum_children sum_q W_{q,c} L_{q'}(xi^E(xi_q)).
// Per axis a child half cb maps fine Lagrange point qd to coarse coordinate (qd + 2 cb - 2)/2.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double lag1(int qp, double x) {
    return qp == 0 ? 0.5 * x * (x - 1.0) : (qp == 1 ? 1.0 - x * x : 0.5 * x * (x + 1.0));
}

template <int DIM>
__global__ void k_coarsen_weights(GridDev gf, GridDev gc, const double *__restrict__ Wf, double *__restrict__ Wc) {
    constexpr int NQ = DIM == 3 ? 27 : 9;
    const int E = blockIdx.x * blockDim.x + threadIdx.x;
    if (E >= gc.nel) return;
    const int I = E % gc.nx, J = (E / gc.nx) % gc.ny, K = E / (gc.nx * gc.ny);
    double T1[2][3][3];  // [cb][qd][q'd]
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int qd = 0; qd < 3; qd++)
#pragma unroll
            for (int qp = 0; qp < 3; qp++) T1[cb][qd][qp] = lag1(qp, 0.5 * (double)(qd + 2 * cb - 2));
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q] = 0.0;
    for (int c = 0; c < (1 << DIM); c++) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const int ef = (2 * I + cx) + gf.nx * ((2 * J + cy) + gf.ny * (DIM == 3 ? (2 * K + cz) : 0));
        for (int q = 0; q < NQ; q++) {
            const double wf = Wf[widx(gf, ef, q)];
            const int qx = q % 3, qy = (q / 3) % 3, qz = q / 9;
#pragma unroll
            for (int qp = 0; qp < NQ; qp++) {
                const int px = qp % 3, py = (qp / 3) % 3, pz = qp / 9;
                double t = T1[cx][qx][px] * T1[cy][qy][py];
                if (DIM == 3) t *= T1[cz][qz][pz];
                acc[qp] += wf * t;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) Wc[widx(gc, E, q)] = acc[q];
}

// restriction b_c = P^T r_f : coarse dof (node (I,J,K), component c) gathers fine nodes 2I+{-1,0,1}
// (weights 1/2, 1, 1/2 per axis) over the free fine dofs.  One thread per coarse dof; the 3^DIM loads are
// unrolled and issued together (out-of-range neighbours read a clamped in-range address with weight 0).
// MASKF false: r_f is already zero on the Dirichlet fine dofs (the residual pass writes it so) and the fine
// mask is not read -- its 27 byte loads per thread cost as much as the values' (3.4% of a 128x128x64 solve).
// With xc set, the coarse level's first Jacobi-Chebyshev step from a zero guess follows in the same pass:
// dc = xc = c2 invdc bc.
template <int DIM, typename T, bool MASKF>
__global__ __launch_bounds__(256) void k_restrict(GridDev gf, GridDev gc, const unsigned char *__restrict__ maskf,
                                                  const unsigned char *__restrict__ maskc, const T *__restrict__ rf,
                                                  T *__restrict__ bc, const T *__restrict__ invdc, T *__restrict__ dc,
                                                  T *__restrict__ xc, T c2) {
    const unsigned ndc = (unsigned)gc.nnx * gc.nny * gc.nnz * DIM;  // 32-bit index arithmetic (< 2^31 dofs)
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ndc) return;
    if (maskc[t]) {
        bc[t] = (T)0;
        if (xc) dc[t] = xc[t] = (T)0;
        return;
    }
    const int c = (int)(t % DIM);
    const unsigned n = t / DIM, nxy = (unsigned)gc.nnx * gc.nny;
    const int K = (int)(n / nxy), rem = (int)(n - (unsigned)K * nxy);
    const int J = rem / gc.nnx, I = rem - J * gc.nnx;
    T acc = (T)0;
#pragma unroll
    for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); dz++) {
        const int k = 2 * K + dz;
        const bool okz = k >= 0 && k < gf.nnz;
        const int kc = okz ? k : 2 * K;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            const int j = 2 * J + dy;
            const bool oky = j >= 0 && j < gf.nny;
            const int jc = oky ? j : 2 * J;
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int i = 2 * I + dx;
                const bool okx = i >= 0 && i < gf.nnx;
                const int ic = okx ? i : 2 * I;
                const unsigned fd = ((unsigned)ic + (unsigned)gf.nnx * ((unsigned)jc + (unsigned)gf.nny * kc)) * DIM + c;
                const T v = rf[fd];
                const bool use = okx && oky && okz && !(MASKF && maskf[fd]);
                const T wt = (T)((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0));
                acc += use ? wt * v : (T)0;
            }
        }
    }
    bc[t] = acc;
    if (xc) {
        const T x0 = c2 * invdc[t] * acc;
        dc[t] = x0;
        xc[t] = x0;
    }
}

// prolongation x_f += P x_c : even index injects, odd index averages its two coarse neighbours
template <int DIM, typename T>
__global__ void k_prolong_add(GridDev gf, GridDev gc, const unsigned char *__restrict__ maskf,
                              const T *__restrict__ xc, T *__restrict__ xf) {
    const int nnf = gf.nnx * gf.nny * gf.nnz;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnf) return;
    const int i = n % gf.nnx, j = (n / gf.nnx) % gf.nny, k = n / (gf.nnx * gf.nny);
    const int ix[2] = {i >> 1, (i >> 1) + (i & 1)}, jy[2] = {j >> 1, (j >> 1) + (j & 1)},
              kz[2] = {k >> 1, (k >> 1) + (k & 1)};
    const int nxo = (i & 1) + 1, nyo = (j & 1) + 1, nzo = (DIM == 3) ? (k & 1) + 1 : 1;
    const T wt = (T)1 / (T)(nxo * nyo * nzo);
    T acc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; c++) acc[c] = (T)0;
    for (int oz = 0; oz < nzo; oz++)
        for (int oy = 0; oy < nyo; oy++)
            for (int ox = 0; ox < nxo; ox++) {
                const size_t cn = (size_t)ix[ox] + gc.nnx * ((size_t)jy[oy] + gc.nny * kz[oz]);
#pragma unroll
                for (int c = 0; c < DIM; c++) acc[c] += xc[cn * DIM + c];
            }
#pragma unroll
    for (int c = 0; c < DIM; c++)
        if (!maskf[(size_t)n * DIM + c]) xf[(size_t)n * DIM + c] += wt * acc[c];
}

// ------------------------------------------------------------------------------------------
// dense coarse operator (FP64): assembled entry by entry, identity on BC dofs
// ------------------------------------------------------------------------------------------
// K_e[(a,i),(b,j)] of element e (vertices a, b; components i, j) from its quadrature weights
template <int DIM>
__device__ __forceinline__ double elem_entry(const GridDev &g, const double *__restrict__ WK, int e, int a, int i,
                                             int b, int j) {
    constexpr int NV = 1 << DIM, NQ = DIM == 3 ? 27 : 9;
    const double s[3] = {2.0 / g.hx, 2.0 / g.hy, 2.0 / g.hz};
    double v = 0.0;
    for (int q = 0; q < NQ; q++) {
        const int qd[3] = {q % 3, (q / 3) % 3, q / 9};
        double ga[3], gb[3];
        for (int k = 0; k < DIM; k++) {
            int ca = sgn((a >> k) & 1), cb = sgn((b >> k) & 1);
            for (int d = 0; d < DIM; d++)
                if (d != k) {
                    ca *= cfac((a >> d) & 1, qd[d]);
                    cb *= cfac((b >> d) & 1, qd[d]);
                }
            ga[k] = s[k] * ca / (double)NV;
            gb[k] = s[k] * cb / (double)NV;
        }
        double m = 0.0;
        if (i == j)
            for (int k = 0; k < DIM; k++) m += ga[k] * gb[k];
        m += ga[j] * gb[i];
        v += WK[widx(g, e, q)] * m;
    }
    return v;
}

// One thread per entry (R, C): the sum over the elements holding both nodes, in a fixed element order, so
// the matrix (and its inverse) is bitwise reproducible.  Rows and columns of Dirichlet dofs are zero here;
// k_dense_bc_identity puts 1 on their diagonal.
template <int DIM>
__global__ void k_dense_assemble(GridDev g, const double *__restrict__ WK, const unsigned char *__restrict__ mask,
                                 double *__restrict__ A, int n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const int R = (int)(t / n), C = (int)(t % n);
    const int na = R / DIM, i = R % DIM, nb = C / DIM, j = C % DIM;
    const int ax = na % g.nnx, ay = (na / g.nnx) % g.nny, az = na / (g.nnx * g.nny);
    const int bx = nb % g.nnx, by = (nb / g.nnx) % g.nny, bz = nb / (g.nnx * g.nny);
    double v = 0.0;
    if (!mask[R] && !mask[C] && abs(ax - bx) <= 1 && abs(ay - by) <= 1 && abs(az - bz) <= 1) {
        // elements (ex, ey, ez) containing both nodes: ex in [max(ax,bx)-1, min(ax,bx)] within the grid
        const int x0 = max(max(ax, bx) - 1, 0), x1 = min(min(ax, bx), g.nx - 1);
        const int y0 = max(max(ay, by) - 1, 0), y1 = min(min(ay, by), g.ny - 1);
        const int z0 = DIM == 3 ? max(max(az, bz) - 1, 0) : 0, z1 = DIM == 3 ? min(min(az, bz), g.nz - 1) : 0;
        for (int ez = z0; ez <= z1; ez++)
            for (int ey = y0; ey <= y1; ey++)
                for (int ex = x0; ex <= x1; ex++) {
                    const int e = ex + g.nx * (ey + g.ny * ez);
                    const int a = (ax - ex) + 2 * (ay - ey) + (DIM == 3 ? 4 * (az - ez) : 0);
                    const int b = (bx - ex) + 2 * (by - ey) + (DIM == 3 ? 4 * (bz - ez) : 0);
                    v += elem_entry<DIM>(g, WK, e, a, i, b, j);
                }
    }
    A[t] = v;
}

__global__ void k_dense_bc_identity(const unsigned char *__restrict__ mask, double *__restrict__ A, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask[i]) A[(size_t)i * n + i] = 1.0;
}

// In-place blocked Gauss-Jordan inverse of an SPD matrix (no pivoting: every pivot block is a Schur complement of
// an SPD matrix, so it is SPD).  Pivot blocks of GJB rows; per block K = [k0, k0 + b), with Pi = A_KK^-1:
//   A_Kj <- Pi A_Kj (j outside K),  A_ij <- A_ij - A_iK A_Kj (i, j outside K, the new A_Kj),
//   A_iK <- -A_iK Pi (i outside K),  A_KK <- Pi;
// after the last block A holds A^-1.  Three launches per block: the pivot block's inverse (one workgroup, the
// scalar Gauss-Jordan steps in LDS), the row panel together with a copy of the column panel (so the update
// reads the column panel as it was), and the update of every row outside K.  Each launch writes only entries
// the launch itself does not read from other threads; the sums run in a fixed order (deterministic).
constexpr int GJB = 32;

__global__ __launch_bounds__(256) void k_gjb_pivot(double *__restrict__ A, int n, int k0, int b,
                                                   double *__restrict__ Pi) {
    __shared__ double M[GJB][GJB + 1];
    const int t = threadIdx.x;
    for (int e = t; e < b * b; e += 256) M[e / b][e % b] = A[(size_t)(k0 + e / b) * n + k0 + e % b];
    __syncthreads();
    for (int p = 0; p < b; p++) {
        // pivot p of the block: new values from the old ones, written after every thread has read
        double v[GJB * GJB / 256];
        const double rp = 1.0 / M[p][p];
#pragma unroll
        for (int q = 0; q < GJB * GJB / 256; q++) {
            const int e = t + 256 * q, r = e / b, c = e % b;
            if (e >= b * b) continue;
            if (r == p) v[q] = c == p ? rp : M[p][c] * rp;
            else if (c == p) v[q] = -M[r][p] * rp;
            else v[q] = M[r][c] - M[r][p] * (M[p][c] * rp);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < GJB * GJB / 256; q++) {
            const int e = t + 256 * q;
            if (e < b * b) M[e / b][e % b] = v[q];
        }
        __syncthreads();
    }
    for (int e = t; e < b * b; e += 256) {
        const int r = e / b, c = e % b;
        A[(size_t)(k0 + r) * n + k0 + c] = M[r][c];
        Pi[r * GJB + c] = M[r][c];
    }
}

// workgroups [0, nrp): the row panel, 64 columns each (A_Kj <- Pi A_Kj, columns outside K); the rest: the copy
// Cw[i][m] = A[i][k0 + m] of the column panel
__global__ __launch_bounds__(256) void k_gjb_panels(double *__restrict__ A, int n, int k0, int b,
                                                    const double *__restrict__ Pi, double *__restrict__ Cw, int nrp) {
    const int t = threadIdx.x;
    if ((int)blockIdx.x < nrp) {
        __shared__ double R[GJB][64];
        __shared__ double Ps[GJB][GJB + 1];
        const int j0 = blockIdx.x * 64;
        for (int e = t; e < b * b; e += 256) Ps[e / b][e % b] = Pi[(e / b) * GJB + e % b];
        for (int e = t; e < b * 64; e += 256) {
            const int r = e / 64, j = j0 + e % 64;
            R[r][e % 64] = j < n ? A[(size_t)(k0 + r) * n + j] : 0.0;
        }
        __syncthreads();
        for (int e = t; e < b * 64; e += 256) {
            const int r = e / 64, c = e % 64, j = j0 + c;
            if (j >= n || (j >= k0 && j < k0 + b)) continue;
            double s = 0.0;
            for (int m = 0; m < b; m++) s += Ps[r][m] * R[m][c];
            A[(size_t)(k0 + r) * n + j] = s;
        }
    } else {
        const size_t e = (size_t)(blockIdx.x - nrp) * 256 + t;
        if (e < (size_t)n * b) {
            const size_t i = e / b;
            const int m = (int)(e % b);
            Cw[i * GJB + m] = A[i * n + k0 + m];
        }
    }
}

// rows outside K: A_ij <- A_ij - Cw_i. A_Kj (j outside K), A_iK <- -Cw_i. Pi; a workgroup holds 16 rows x 64
// columns (4 rows per thread) and stages the 64 columns' rows of K (or of Pi) and its rows' column panel
constexpr int GJR = 16;
__global__ __launch_bounds__(256) void k_gjb_update(double *__restrict__ A, int n, int k0, int b,
                                                    const double *__restrict__ Pi, const double *__restrict__ Cw) {
    __shared__ double R[GJB][64];
    __shared__ double Cs[GJR][GJB + 1];
    const int t = threadIdx.x, j0 = blockIdx.x * 64, i0 = blockIdx.y * GJR;
    for (int e = t; e < b * 64; e += 256) {
        const int m = e / 64, j = j0 + e % 64;
        R[m][e % 64] = j >= n ? 0.0 : (j >= k0 && j < k0 + b) ? Pi[m * GJB + (j - k0)] : A[(size_t)(k0 + m) * n + j];
    }
    for (int e = t; e < GJR * GJB; e += 256) {
        const int r = e / GJB, m = e % GJB, i = i0 + r;
        Cs[r][m] = (i < n && m < b) ? Cw[(size_t)i * GJB + m] : 0.0;
    }
    __syncthreads();
    const int c = t & 63, j = j0 + c;
    if (j >= n) return;
    const bool inK = j >= k0 && j < k0 + b;
#pragma unroll
    for (int q = 0; q < GJR / 4; q++) {
        const int r = (t >> 6) + 4 * q, i = i0 + r;
        if (i >= n || (i >= k0 && i < k0 + b)) continue;
        double s = inK ? 0.0 : A[(size_t)i * n + j];
        for (int m = 0; m < b; m++) s -= Cs[r][m] * R[m][c];
        A[(size_t)i * n + j] = s;
    }
}

// y = Ainv x: one workgroup (4 wavefronts) per row
template <typename T>
__global__ __launch_bounds__(256) void k_dense_matvec(const double *__restrict__ Ainv, int n,
                                                      const T *__restrict__ x, T *__restrict__ y) {
    __shared__ double part[4];
    const int row = blockIdx.x;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) s += Ainv[(size_t)row * n + j] * (double)x[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) y[row] = (T)((part[0] + part[1]) + (part[2] + part[3]));
}

// ------------------------------------------------------------------------------------------
extern "C" {

// Element-per-thread variant with per-dof atomics: the plain restatement the tiled kernels of
// ust_apply.hip are checked against.
int ust_k_apply_K_ref_f64(const ust_grid *gr, const double *WK, const unsigned char *mask, const double *u, double *y,
                          void *stream) {
    GridDev g = grid_dev(gr);
    if (!WK || !mask || !u || !y) return UST_EINVAL;
    const size_t ndof = (size_t)g.nnx * g.nny * g.nnz * g.dim;
    hipStream_t st = ust_stream(stream);
    hipLaunchKernelGGL(k_fill<double>, dim3(ust_grid_size(ndof, UST_BLOCK)), dim3(UST_BLOCK), 0, st, y, 0.0, ndof);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_apply_K<3, double, 0, 1>), dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           st, g, WK, mask, u, y);
    else
        hipLaunchKernelGGL((k_apply_K<2, double, 0, 1>), dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           st, g, WK, mask, u, y);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_apply_K_ref_f32(const ust_grid *gr, const float *WK, const unsigned char *mask, const float *u, float *y,
                          void *stream) {
    GridDev g = grid_dev(gr);
    if (!WK || !mask || !u || !y) return UST_EINVAL;
    const size_t ndof = (size_t)g.nnx * g.nny * g.nnz * g.dim;
    hipStream_t st = ust_stream(stream);
    hipLaunchKernelGGL(k_fill<float>, dim3(ust_grid_size(ndof, UST_BLOCK)), dim3(UST_BLOCK), 0, st, y, 0.0f, ndof);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_apply_K<3, float, 0, 1>), dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           st, g, WK, mask, u, y);
    else
        hipLaunchKernelGGL((k_apply_K<2, float, 0, 1>), dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           st, g, WK, mask, u, y);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_lift_K(const ust_grid *gr, const double *WK, const unsigned char *mask, const double *ubc, double *y,
                 void *stream) {
    GridDev g = grid_dev(gr);
    if (!WK || !mask || !ubc || !y) return UST_EINVAL;
    if (g.dim == 3)
        hipLaunchKernelGGL((k_apply_K<3, double, 1, -1>), dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WK, mask, ubc, y);
    else
        hipLaunchKernelGGL((k_apply_K<2, double, 1, -1>), dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0,
                           ust_stream(stream), g, WK, mask, ubc, y);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_apply_G(const ust_grid *gr, const double *WG, const unsigned char *mask, const double *p, double *y,
                  void *stream) {
    GridDev g = grid_dev(gr);
    if (!WG || !mask || !p || !y) return UST_EINVAL;
    const size_t ndof = (size_t)g.nnx * g.nny * g.nnz * g.dim;
    hipStream_t st = ust_stream(stream);
    const size_t nnodes = ndof / g.dim;
    if (g.dim == 3)
        hipLaunchKernelGGL((k_apply_G<3, false>), dim3(ust_grid_size(nnodes, UST_BLOCK)), dim3(UST_BLOCK), 0, st, g, WG, mask,
                           p, y);
    else
        hipLaunchKernelGGL((k_apply_G<2, false>), dim3(ust_grid_size(nnodes, UST_BLOCK)), dim3(UST_BLOCK), 0, st, g, WG, mask,
                           p, y);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_abs_G_rowsum(const ust_grid *gr, const double *WG, const unsigned char *mask, double *y, void *stream) {
    GridDev g = grid_dev(gr);
    if (!WG || !mask || !y) return UST_EINVAL;
    const unsigned grid = ust_grid_size((size_t)g.nnx * g.nny * g.nnz, UST_BLOCK);
    if (g.dim == 3)
        hipLaunchKernelGGL((k_apply_G<3, true>), dim3(grid), dim3(UST_BLOCK), 0, ust_stream(stream), g, WG, mask,
                           (const double *)nullptr, y);
    else
        hipLaunchKernelGGL((k_apply_G<2, true>), dim3(grid), dim3(UST_BLOCK), 0, ust_stream(stream), g, WG, mask,
                           (const double *)nullptr, y);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_apply_Gt(const ust_grid *gr, const double *WG, const unsigned char *mask, const double *u, double *q,
                   int bc_only, void *stream) {
    GridDev g = grid_dev(gr);
    if (!WG || !mask || !u || !q) return UST_EINVAL;
    hipStream_t st = ust_stream(stream);
    const unsigned grid = ust_grid_size(g.nel, UST_BLOCK);
    if (g.dim == 3) {
        if (bc_only == 1) hipLaunchKernelGGL((k_apply_Gt<3, 1>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
        else if (bc_only == 2) hipLaunchKernelGGL((k_apply_Gt<3, 2>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
        else if (bc_only == 3) hipLaunchKernelGGL((k_apply_Gt<3, 3>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
        else hipLaunchKernelGGL((k_apply_Gt<3, 0>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
    } else {
        if (bc_only == 1) hipLaunchKernelGGL((k_apply_Gt<2, 1>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
        else if (bc_only == 2) hipLaunchKernelGGL((k_apply_Gt<2, 2>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
        else if (bc_only == 3) hipLaunchKernelGGL((k_apply_Gt<2, 3>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
        else hipLaunchKernelGGL((k_apply_Gt<2, 0>), dim3(grid), dim3(UST_BLOCK), 0, st, g, WG, mask, u, q);
    }
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_force_dir(const ust_grid *gr, const double *WF, double gravity, const double *dir,
                    const unsigned char *mask, double *f, void *stream) {
    GridDev g = grid_dev(gr);
    if (!WF || !mask || !f || !dir) return UST_EINVAL;
    Dir3 d = {{dir[0], dir[1], g.dim == 3 ? dir[2] : 0.0}};
    const unsigned grid = ust_grid_size((size_t)g.nnx * g.nny * g.nnz, UST_BLOCK);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_force<3>, dim3(grid), dim3(UST_BLOCK), 0, ust_stream(stream), g, WF, gravity, d, mask, f);
    else
        hipLaunchKernelGGL(k_force<2>, dim3(grid), dim3(UST_BLOCK), 0, ust_stream(stream), g, WF, gravity, d, mask, f);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_force(const ust_grid *gr, const double *WF, double gravity, const unsigned char *mask, double *f,
                void *stream) {
    const double down[3] = {0.0, -1.0, 0.0};
    return ust_k_force_dir(gr, WF, gravity, down, mask, f, stream);
}

int ust_k_inv_diag(const ust_grid *gr, const double *WK, const unsigned char *mask, double *invd64, float *invd32,
                   void *stream) {
    GridDev g = grid_dev(gr);
    if (!WK || !mask || !invd64) return UST_EINVAL;
    const size_t ndof = (size_t)g.nnx * g.nny * g.nnz * g.dim;
    hipStream_t st = ust_stream(stream);
    hipLaunchKernelGGL(k_fill<double>, dim3(ust_grid_size(ndof, UST_BLOCK)), dim3(UST_BLOCK), 0, st, invd64, 0.0,
                       ndof);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_diag_acc<3>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0, st, g, WK, mask,
                           invd64);
    else
        hipLaunchKernelGGL(k_diag_acc<2>, dim3(ust_grid_size(g.nel, UST_BLOCK)), dim3(UST_BLOCK), 0, st, g, WK, mask,
                           invd64);
    hipLaunchKernelGGL(k_diag_invert, dim3(ust_grid_size(ndof, UST_BLOCK)), dim3(UST_BLOCK), 0, st, mask, invd64,
                       invd32, ndof);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

static int coarse_grid(const ust_grid *gf, ust_grid *gc) {
    *gc = *gf;
    for (int d = 0; d < gf->dim; d++) {
        if (gf->nel[d] % 2) return UST_EINVAL;
        gc->nel[d] = gf->nel[d] / 2;
        gc->h[d] = gf->h[d] * 2.0;
    }
    return UST_OK;
}

int ust_k_coarsen_weights(const ust_grid *gf, const double *WKf, double *WKc, void *stream) {
    ust_grid gcr;
    if (coarse_grid(gf, &gcr) != UST_OK || !WKf || !WKc) return UST_EINVAL;
    GridDev f = grid_dev(gf), c = grid_dev(&gcr);
    if (f.dim == 3)
        hipLaunchKernelGGL(k_coarsen_weights<3>, dim3(ust_grid_size(c.nel, 128)), dim3(128), 0, ust_stream(stream), f,
                           c, WKf, WKc);
    else
        hipLaunchKernelGGL(k_coarsen_weights<2>, dim3(ust_grid_size(c.nel, 128)), dim3(128), 0, ust_stream(stream), f,
                           c, WKf, WKc);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

#define TRANSFER_IMPL(T, SUF)                                                                                    \
    static int restrict_##SUF(const ust_grid *gf, const unsigned char *maskf, const unsigned char *maskc,      \
                              const T *rf, T *bc, const T *invdc, T *dc, T *xc, T c2, void *stream) {         \
        ust_grid gcr;                                                                                          \
        if (coarse_grid(gf, &gcr) != UST_OK || !maskc || !rf || !bc) return UST_EINVAL;                        \
        GridDev f = grid_dev(gf), c = grid_dev(&gcr);                                                          \
        const int nnc = c.nnx * c.nny * c.nnz;                                                                 \
        const dim3 grid(ust_grid_size((size_t)nnc * f.dim, UST_BLOCK));                                        \
        hipStream_t st = ust_stream(stream);                                                                   \
        if (f.dim == 3 && maskf)                                                                               \
            hipLaunchKernelGGL((k_restrict<3, T, true>), grid, dim3(UST_BLOCK), 0, st, f, c, maskf, maskc, rf, \
                               bc, invdc, dc, xc, c2);                                                         \
        else if (f.dim == 3)                                                                                   \
            hipLaunchKernelGGL((k_restrict<3, T, false>), grid, dim3(UST_BLOCK), 0, st, f, c, maskf, maskc,    \
                               rf, bc, invdc, dc, xc, c2);                                                     \
        else if (maskf)                                                                                        \
            hipLaunchKernelGGL((k_restrict<2, T, true>), grid, dim3(UST_BLOCK), 0, st, f, c, maskf, maskc, rf, \
                               bc, invdc, dc, xc, c2);                                                         \
        else                                                                                                   \
            hipLaunchKernelGGL((k_restrict<2, T, false>), grid, dim3(UST_BLOCK), 0, st, f, c, maskf, maskc,    \
                               rf, bc, invdc, dc, xc, c2);                                                     \
        UST_LAUNCH_CHECK();                                                                                    \
        return UST_OK;                                                                                         \
    }                                                                                                          \
    int ust_k_restrict_##SUF(const ust_grid *gf, const unsigned char *maskf, const unsigned char *maskc,       \
                             const T *rf, T *bc, void *stream) {                                               \
        return restrict_##SUF(gf, maskf, maskc, rf, bc, nullptr, nullptr, nullptr, (T)0, stream);              \
    }                                                                                                          \
    int ust_k_restrict_cheby_first_##SUF(const ust_grid *gf, const unsigned char *maskf,                       \
                                         const unsigned char *maskc, const T *rf, T *bc, const T *invdc,       \
                                         T *dc, T *xc, T c2, void *stream) {                                   \
        if (!invdc || !dc || !xc) return UST_EINVAL;                                                           \
        return restrict_##SUF(gf, maskf, maskc, rf, bc, invdc, dc, xc, c2, stream);                            \
    }                                                                                                          \
    int ust_k_prolong_add_##SUF(const ust_grid *gf, const unsigned char *maskf, const T *xc, T *xf,           \
                                void *stream) {                                                                \
        ust_grid gcr;                                                                                          \
        if (coarse_grid(gf, &gcr) != UST_OK || !maskf || !xc || !xf) return UST_EINVAL;                        \
        GridDev f = grid_dev(gf), c = grid_dev(&gcr);                                                          \
        const int nnf = f.nnx * f.nny * f.nnz;                                                                 \
        if (f.dim == 3)                                                                                        \
            hipLaunchKernelGGL((k_prolong_add<3, T>), dim3(ust_grid_size(nnf, UST_BLOCK)), dim3(UST_BLOCK), 0, \
                               ust_stream(stream), f, c, maskf, xc, xf);                                       \
        else                                                                                                   \
            hipLaunchKernelGGL((k_prolong_add<2, T>), dim3(ust_grid_size(nnf, UST_BLOCK)), dim3(UST_BLOCK), 0, \
                               ust_stream(stream), f, c, maskf, xc, xf);                                       \
        UST_LAUNCH_CHECK();                                                                                    \
        return UST_OK;                                                                                         \
    }
TRANSFER_IMPL(double, f64)
TRANSFER_IMPL(float, f32)

int ust_k_dense_assemble(const ust_grid *gr, const double *WK, const unsigned char *mask, double *A, void *stream) {
    GridDev g = grid_dev(gr);
    const int n = g.nnx * g.nny * g.nnz * g.dim;
    if (!WK || !mask || !A || n > 4096) return UST_EINVAL;
    hipStream_t st = ust_stream(stream);
    const size_t work = (size_t)n * n;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_dense_assemble<3>, dim3(ust_grid_size(work, UST_BLOCK)), dim3(UST_BLOCK), 0, st, g, WK,
                           mask, A, n);
    else
        hipLaunchKernelGGL(k_dense_assemble<2>, dim3(ust_grid_size(work, UST_BLOCK)), dim3(UST_BLOCK), 0, st, g, WK,
                           mask, A, n);
    hipLaunchKernelGGL(k_dense_bc_identity, dim3(ust_grid_size(n, UST_BLOCK)), dim3(UST_BLOCK), 0, st, mask, A, n);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_dense_spd_inverse(double *A, int n, double *work, void *stream) {
    if (!A || !work || n <= 0 || n > 4096) return UST_EINVAL;
    hipStream_t st = ust_stream(stream);
    double *Pi = work, *Cw = work + GJB * GJB;
    const int nrp = (n + 63) / 64;
    const dim3 ugrid((unsigned)nrp, (unsigned)((n + GJR - 1) / GJR));
    for (int k0 = 0; k0 < n; k0 += GJB) {
        const int b = n - k0 < GJB ? n - k0 : GJB;
        const unsigned ncp = (unsigned)(((size_t)n * b + 255) / 256);
        hipLaunchKernelGGL(k_gjb_pivot, dim3(1), dim3(256), 0, st, A, n, k0, b, Pi);
        hipLaunchKernelGGL(k_gjb_panels, dim3((unsigned)nrp + ncp), dim3(256), 0, st, A, n, k0, b, Pi, Cw, nrp);
        hipLaunchKernelGGL(k_gjb_update, ugrid, dim3(256), 0, st, A, n, k0, b, Pi, Cw);
    }
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_dense_matvec_f64(const double *Ainv, int n, const double *x, double *y, void *stream) {
    if (!Ainv || !x || !y || n <= 0) return UST_EINVAL;
    hipLaunchKernelGGL(k_dense_matvec<double>, dim3(n), dim3(256), 0, ust_stream(stream), Ainv, n, x, y);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

int ust_k_dense_matvec_f32(const double *Ainv, int n, const float *x, float *y, void *stream) {
    if (!Ainv || !x || !y || n <= 0) return UST_EINVAL;
    hipLaunchKernelGGL(k_dense_matvec<float>, dim3(n), dim3(256), 0, ust_stream(stream), Ainv, n, x, y);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

}  // extern "C"
