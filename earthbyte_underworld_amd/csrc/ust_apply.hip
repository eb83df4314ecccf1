// The hot operator of the BSSCR solve on gfx950: y = K u, matrix-free, from the 3^dim quadrature
// weights of each element (the MatMult of the A11 block, auglag-driver-DGTGD.c:191-403, that every
// inner Krylov iteration and every multigrid smoothing step performs, PETScMGSolver.c:641-716).
//
// 3D: one 256-thread workgroup owns a 16x16 column of elements, one element per thread, and streams a
// chunk of element layers along z.  Node planes are staged in LDS once each; per layer every thread applies
// its element with sum factorisation (gradients at the Lagrange points from edge differences, stress, and
// the transposed interpolation back to the vertices, ~0.8 kflop per element) and leaves its 8 vertex
// results in LDS; each node of the plane below the layer then sums, in a fixed order, the contributions of
// the (up to) four elements above it and the four below it (carried in registers from the previous layer),
// and is finished: written to y through the epilogue, unless it lies on the ring of the column (shared with
// a neighbouring column) or on a chunk-boundary plane (shared with the next chunk).  Those nodes go to the
// workgroup's slot of a workspace, and a second, node-ordered pass sums the two to eight slots covering each
// of them.  No atomics touch memory and the summation order is fixed, so K u is bitwise reproducible.
// Grids with fewer than 1024 column layers (the coarse multigrid levels) use 16x4x4 tiles instead, one per
// workgroup, which gives small grids more workgroups.
// 2D: one workgroup per 16x16 tile, node block reduced in LDS, the same workspace scheme on the tile faces.
#include <mutex>
#include <stdio.h>

#include "ust_common.h"

// axes orthogonal to k, in increasing order
__host__ __device__ constexpr int P1(int k) { return k == 0 ? 1 : 0; }
__host__ __device__ constexpr int P2(int k) { return k == 2 ? 1 : 2; }
__host__ __device__ constexpr int bit_on(int axis, int k, int bk, int b1, int b2) {
    return axis == k ? bk : (axis == P1(k) ? b1 : b2);
}
// vertex index a = bx + 2 by + 4 bz with bit bk on axis k and (b1, b2) on the orthogonal axes
__host__ __device__ constexpr int node3(int k, int bk, int b1, int b2) {
    return bit_on(0, k, bk, b1, b2) + 2 * bit_on(1, k, bk, b1, b2) + 4 * bit_on(2, k, bk, b1, b2);
}

// 2x2 vertex values -> 3x3 Lagrange points {-1, 0, 1}^2 (bilinear interpolation)
template <typename T>
__device__ __forceinline__ void interp22(const T D[2][2], T G[3][3]) {
    const T h = (T)0.5;
    G[0][0] = D[0][0];
    G[0][2] = D[0][1];
    G[2][0] = D[1][0];
    G[2][2] = D[1][1];
    G[1][0] = h * (D[0][0] + D[1][0]);
    G[1][2] = h * (D[0][1] + D[1][1]);
    G[0][1] = h * (D[0][0] + D[0][1]);
    G[2][1] = h * (D[1][0] + D[1][1]);
    G[1][1] = (T)0.25 * (D[0][0] + D[0][1] + D[1][0] + D[1][1]);
}

// transpose of interp22
template <typename T>
__device__ __forceinline__ void interp22T(const T R[3][3], T O[2][2]) {
    const T h = (T)0.5, c = (T)0.25 * R[1][1];
    O[0][0] = R[0][0] + h * (R[1][0] + R[0][1]) + c;
    O[1][0] = R[2][0] + h * (R[1][0] + R[2][1]) + c;
    O[0][1] = R[0][2] + h * (R[0][1] + R[1][2]) + c;
    O[1][1] = R[2][2] + h * (R[2][1] + R[1][2]) + c;
}

// physical gradient d u_i / d x_k at the 9 points orthogonal to k (constant along k):
// gp[q1][q2] = (s_k / 2) * interp22(u(k=1) - u(k=0))
template <typename T, int I, int K, typename U>
__device__ __forceinline__ void grad3(const U &u, T sk, T gp[3][3]) {
    T D[2][2];
#pragma unroll
    for (int b1 = 0; b1 < 2; b1++)
#pragma unroll
        for (int b2 = 0; b2 < 2; b2++)
            D[b1][b2] = (T)0.5 * sk * (u(node3(K, 1, b1, b2), I) - u(node3(K, 0, b1, b2), I));
    interp22<T>(D, gp);
}

// y[.][I] += (s_K/2) sgn(a_K) interp22T(R)  for derivative direction K
template <typename T, int I, int K>
__device__ __forceinline__ void scatter3(const T R[3][3], T sk, T y[8][3]) {
    T O[2][2];
    interp22T<T>(R, O);
#pragma unroll
    for (int b1 = 0; b1 < 2; b1++)
#pragma unroll
        for (int b2 = 0; b2 < 2; b2++) {
            const T v = (T)0.5 * sk * O[b1][b2];
            y[node3(K, 1, b1, b2)][I] += v;
            y[node3(K, 0, b1, b2)][I] -= v;
        }
}

template <typename T, int I, typename U>
__device__ __forceinline__ void diag3(const U &u, const T W[27], const T s[3], T y[8][3]) {
    T gp[3][3], R[3][3];
    grad3<T, I, I>(u, s[I], gp);
#pragma unroll
    for (int q1 = 0; q1 < 3; q1++)
#pragma unroll
        for (int q2 = 0; q2 < 3; q2++) {
            T ws = (T)0;
#pragma unroll
            for (int qk = 0; qk < 3; qk++) {
                int qs[3];
                qs[I] = qk;
                qs[P1(I)] = q1;
                qs[P2(I)] = q2;
                ws += W[qs[0] + 3 * qs[1] + 9 * qs[2]];
            }
            R[q1][q2] = (T)2 * gp[q1][q2] * ws;
        }
    scatter3<T, I, I>(R, s[I], y);
}

template <typename T, int I, int K, typename U>
__device__ __forceinline__ void pair3(const U &u, const T W[27], const T s[3], T y[8][3]) {
    T gik[3][3], gki[3][3], Rik[3][3], Rki[3][3];
    grad3<T, I, K>(u, s[K], gik);  // d u_I / d x_K : indexed by the axes orthogonal to K
    grad3<T, K, I>(u, s[I], gki);  // d u_K / d x_I : indexed by the axes orthogonal to I
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            Rik[a][b] = (T)0;
            Rki[a][b] = (T)0;
        }
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const int qs[3] = {q % 3, (q / 3) % 3, q / 9};
        const T t = W[q] * (gik[qs[P1(K)]][qs[P2(K)]] + gki[qs[P1(I)]][qs[P2(I)]]);
        Rik[qs[P1(K)]][qs[P2(K)]] += t;
        Rki[qs[P1(I)]][qs[P2(I)]] += t;
    }
    scatter3<T, I, K>(Rik, s[K], y);
    scatter3<T, K, I>(Rki, s[I], y);
}

// u(a, i): component i of the element's vertex a (read where needed, so that the velocities need not all be
// held in registers at once)
template <typename T, typename U>
__device__ __forceinline__ void element_apply3(const U &u, const T W[27], const T s[3], T y[8][3]) {
    diag3<T, 0>(u, W, s, y);
    diag3<T, 1>(u, W, s, y);
    diag3<T, 2>(u, W, s, y);
    pair3<T, 0, 1>(u, W, s, y);
    pair3<T, 0, 2>(u, W, s, y);
    pair3<T, 1, 2>(u, W, s, y);
}

// ---------------------------------------- 2D ----------------------------------------------
// vertex a = bx + 2 by; direction k, orthogonal bit b
__host__ __device__ constexpr int node2(int k, int bk, int b) { return k == 0 ? bk + 2 * b : b + 2 * bk; }

template <typename T, int I, int K>
__device__ __forceinline__ void grad2(const T u[4][2], T sk, T gp[3]) {
    const T d0 = (T)0.5 * sk * (u[node2(K, 1, 0)][I] - u[node2(K, 0, 0)][I]);
    const T d1 = (T)0.5 * sk * (u[node2(K, 1, 1)][I] - u[node2(K, 0, 1)][I]);
    gp[0] = d0;
    gp[1] = (T)0.5 * (d0 + d1);
    gp[2] = d1;
}

template <typename T, int I, int K>
__device__ __forceinline__ void scatter2(const T R[3], T sk, T y[4][2]) {
    const T o0 = R[0] + (T)0.5 * R[1], o1 = R[2] + (T)0.5 * R[1];
    const T v0 = (T)0.5 * sk * o0, v1 = (T)0.5 * sk * o1;
    y[node2(K, 1, 0)][I] += v0;
    y[node2(K, 0, 0)][I] -= v0;
    y[node2(K, 1, 1)][I] += v1;
    y[node2(K, 0, 1)][I] -= v1;
}

template <typename T>
__device__ __forceinline__ void element_apply2(const T u[4][2], const T W[9], const T s[3], T y[4][2]) {
    // diagonal terms: R_ii[q_perp] = 2 gp_ii * sum over q_i of W
    {
        T gp[3], R[3];
        grad2<T, 0, 0>(u, s[0], gp);
#pragma unroll
        for (int q = 0; q < 3; q++) R[q] = (T)2 * gp[q] * (W[0 + 3 * q] + W[1 + 3 * q] + W[2 + 3 * q]);
        scatter2<T, 0, 0>(R, s[0], y);
        grad2<T, 1, 1>(u, s[1], gp);
#pragma unroll
        for (int q = 0; q < 3; q++) R[q] = (T)2 * gp[q] * (W[q] + W[q + 3] + W[q + 6]);
        scatter2<T, 1, 1>(R, s[1], y);
    }
    // shear pair (0,1): tau = W (du0/dx1 + du1/dx0); R_01 sums over q1 (indexed q0), R_10 over q0
    {
        T g01[3], g10[3], R01[3] = {0, 0, 0}, R10[3] = {0, 0, 0};
        grad2<T, 0, 1>(u, s[1], g01);  // indexed by q0
        grad2<T, 1, 0>(u, s[0], g10);  // indexed by q1
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int q0 = q % 3, q1 = q / 3;
            const T t = W[q] * (g01[q0] + g10[q1]);
            R01[q0] += t;
            R10[q1] += t;
        }
        scatter2<T, 0, 1>(R01, s[1], y);
        scatter2<T, 1, 0>(R10, s[0], y);
    }
}

// ------------------------------------------------------------------------------------------
// Phase 1: one workgroup per tile.  The tile's nodal velocities are staged in LDS, each thread
// applies its element, and the element results are reduced into the tile's node block in LDS.
// Nodes that belong to this tile only are written to y directly; nodes on a face shared with a
// neighbouring tile go to the tile's slot of a workspace.  Phase 2: one thread per shared dof sums
// the 2, 4 or 8 slots that hold it.  No atomics touch global memory and the summation order is
// fixed, so K u is bitwise reproducible.
// ------------------------------------------------------------------------------------------

// does node coordinate gc lie on a face between two tiles along an axis of tile size T (tn tiles)?
__device__ __forceinline__ bool on_tile_face(int gc, int T, int tn) {
    return gc > 0 && gc % T == 0 && gc / T < tn;
}

// What an apply writes for dof i whose product (K u)_i is v (v = 0 on Dirichlet dofs):
//   EPI_Y       y_i = v
//   EPI_RES     y_i = b_i - v, 0 on Dirichlet dofs             (multigrid residual on the free dofs)
//   EPI_CHEBY   r = b_i - v; d_i = c1 d_i + c2 invd_i r; x_i += d_i with x the input u: one Jacobi-
//               Chebyshev smoothing step in place.  In place is safe: a node interior to a tile is
//               staged only by that tile, before its update, and tile-face nodes are updated in
//               phase 2, after every tile has staged them.
//   EPI_CHEBY0  the same with c1 = 0 (the previous d is not read)
//   EPI_PCHEBY0 (3D) the multigrid prolongation x += P xc of the coarse-grid correction xc, then the
//               EPI_CHEBY0 step on the corrected x, in one pass: the staged velocities are x + P xc,
//               interpolated from the tile's coarse node block held in LDS.
//   EPI_RES0    the first Jacobi-Chebyshev step from a zero guess, x = c2 invd b, then the residual
//               r = b - K x (0 on Dirichlet dofs), in one pass: the staged velocities are c2 invd b (u is not
//               read); y = r, d = x (the previous d is neither read nor written).
enum { EPI_Y = 0, EPI_RES = 1, EPI_CHEBY = 2, EPI_CHEBY0 = 3, EPI_PCHEBY0 = 4, EPI_RES0 = 5 };
template <typename T>
struct Epi {
    T *y;  // EPI_Y / EPI_RES / EPI_RES0: output; EPI_CHEBY*: x after the step
    const T *b;
    const T *invd;
    T *d;  // EPI_CHEBY*: the Chebyshev direction; EPI_RES0: x after the first step
    T c1, c2;
    const T *xc;  // EPI_PCHEBY0: coarse-grid correction (nodes of the grid with half the elements per axis)
    const T *xs;  // EPI_CHEBY*: x before the step, the input u (y itself when the step runs in place)
};

__host__ __device__ constexpr bool cheby_mode(int m) { return m == EPI_CHEBY || m == EPI_CHEBY0 || m == EPI_PCHEBY0; }

// (P xc) at fine node (i, j, k): per axis the coarse node i/2 (even i) or the mean of (i-1)/2 and (i+1)/2
// (odd i), trilinear over the axes -- the weights of k_prolong_add.  Branch-free: always 8 terms, the
// second node of an even axis carries weight 0.  cval(I, J, K) returns a coarse value.
template <typename T, typename F>
__device__ __forceinline__ T prolong_value(int i, int j, int k, F cval) {
    const int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
    const int I1 = I0 + (i & 1), J1 = J0 + (j & 1), K1 = K0 + (k & 1);
    const T wx = (i & 1) ? (T)0.5 : (T)1, wy = (j & 1) ? (T)0.5 : (T)1, wz = (k & 1) ? (T)0.5 : (T)1;
    const T ox = (i & 1) ? (T)0.5 : (T)0, oy = (j & 1) ? (T)0.5 : (T)0, oz = (k & 1) ? (T)0.5 : (T)0;
    const T a0 = wx * cval(I0, J0, K0) + ox * cval(I1, J0, K0);
    const T a1 = wx * cval(I0, J1, K0) + ox * cval(I1, J1, K0);
    const T a2 = wx * cval(I0, J0, K1) + ox * cval(I1, J0, K1);
    const T a3 = wx * cval(I0, J1, K1) + ox * cval(I1, J1, K1);
    return wz * (wy * a0 + oy * a1) + oz * (wy * a2 + oy * a3);
}

// the velocity a pass stages for dof i: u, or for EPI_RES0 the first Chebyshev step from zero, c2 invd b
template <int MODE, typename T>
__device__ __forceinline__ T stage_load(const Epi<T> &ep, const T *u, unsigned i) {
    if (MODE == EPI_RES0) return ep.c2 * ep.invd[i] * ep.b[i];
    return u[i];
}

// operands the epilogue reads, loaded ahead of the stores so that a thread's loads overlap
template <typename T>
struct EpiOps {
    unsigned char m;
    T b, invd, d, x;
};

template <int MODE, typename T>
__device__ __forceinline__ EpiOps<T> epilogue_load(const Epi<T> &ep, const unsigned char *mask, unsigned i) {
    EpiOps<T> o;
    o.m = mask[i];
    o.b = o.invd = o.d = o.x = (T)0;
    if (MODE != EPI_Y) o.b = ep.b[i];
    if (MODE == EPI_RES0) o.invd = ep.invd[i];
    if (cheby_mode(MODE)) {
        o.invd = ep.invd[i];
        o.x = ep.xs[i];
    }
    if (MODE == EPI_CHEBY) o.d = ep.d[i];
    return o;
}

// v: the product (K u)_i before masking
template <int MODE, typename T>
__device__ __forceinline__ void epilogue_store(const Epi<T> &ep, const EpiOps<T> &o, unsigned i, T v) {
    v = o.m ? (T)0 : v;
    if (MODE == EPI_Y) {
        ep.y[i] = v;
    } else if (MODE == EPI_RES || MODE == EPI_RES0) {
        ep.y[i] = o.m ? (T)0 : o.b - v;
        if (MODE == EPI_RES0) ep.d[i] = ep.c2 * o.invd * o.b;  // as staged (kv_cheby_first's rounding)
    } else {
        const T dn = (MODE == EPI_CHEBY ? ep.c1 * o.d : (T)0) + ep.c2 * o.invd * (o.b - v);
        if (MODE == EPI_CHEBY || ep.d) ep.d[i] = dn;  // a first step may leave d unwritten (no later step reads it)
        ep.y[i] = o.x + dn;
    }
}

// ------------------------------------------------------------------------------------------
// 3D column kernel.  Node plane of a 16x16 column: 17x17 nodes, 867 dofs, NR = 4 per thread.
// ------------------------------------------------------------------------------------------
constexpr int CPX = UST_TX3 + 1, CPY = UST_TY3 + 1, CNP = CPX * CPY, CNPD = CNP * 3;
constexpr int CNR = (CNPD + UST_TILE - 1) / UST_TILE;
constexpr int EBS = UST_TILE + 1;  // row stride of the element-result buffer (rows shift across LDS banks)

// Build-time tuning of the column passes.  UST_COL_MINB_F32: workgroups per CU the register budget of the FP32
// passes must allow.  UST_COL_PF_*: when the next layer's node plane and weights are loaded -- 1: at the start of
// the layer, held in registers across the element apply; 2: right after the element apply, so that they are live
// only across the reduction and the staging; 0: when they are used.  Measured on the 128x128x64 passes (MI355X,
// 50 launches each): loads after the element apply take the FP32 residual pass from 58.2 to 55.0 us (168
// instead of 246 VGPRs), the prolongation pass from 75.3 to 72.1 us and the FP64 product from 92.0 to 90.7 us
// (its loads otherwise wait at the layer); 4 FP32 workgroups per CU without prefetch speed the residual pass
// (50 us) but spill in the prolongation pass (112 us), 3 spill in both.
#ifndef UST_COL_MINB_F32
#define UST_COL_MINB_F32 2
#endif
#ifndef UST_COL_PF_F32
#define UST_COL_PF_F32 2
#endif
#ifndef UST_COL_PF_F64
#define UST_COL_PF_F64 2
#endif
// The FP32 product and residual passes (EPI_Y, EPI_RES) take their own values: four workgroups per CU, loads
// when used, on a grid of ~1024 workgroups (UST_COL_WG_TARGET_F32_YR).  Measured on the 128x128x64 passes:
// 53.5 -> 49.5 us (product) and 55.2 -> 51.8 us (residual) against the settings of the other passes; the same
// register budget on the 512-workgroup grid was slower (61 us), and the 1024-workgroup grid slows the
// prolongation pass (71.5 -> 77.9 us) and the FP64 product (88.8 -> 96.1 us), which keep 512.
#ifndef UST_COL_MINB_F32_YR
#define UST_COL_MINB_F32_YR 4
#endif
#ifndef UST_COL_PF_F32_YR
#define UST_COL_PF_F32_YR 0
#endif
__host__ __device__ constexpr int col_minb(int tsize, int mode) {
    return tsize == 4 ? (mode == EPI_Y || mode == EPI_RES || mode == EPI_RES0 ? UST_COL_MINB_F32_YR : UST_COL_MINB_F32)
                      : 2;
}
__host__ __device__ constexpr int col_pf(int tsize, int mode) {
    return tsize == 4 ? (mode == EPI_Y || mode == EPI_RES || mode == EPI_RES0 ? UST_COL_PF_F32_YR : UST_COL_PF_F32)
                      : UST_COL_PF_F64;
}
constexpr long long col_wg_target(int tsize, int mode) {
    return tsize == 4 && (mode == EPI_Y || mode == EPI_RES || mode == EPI_RES0) ? UST_COL_WG_TARGET_F32_YR
                                                                              : UST_COL_WG_TARGET;
}
template <typename T, typename WT, int MODE>
__global__ __launch_bounds__(256, col_minb(sizeof(T), MODE)) void k_applyK3_col(GridDev g, const WT *__restrict__ W,
                                                     const unsigned char *__restrict__ mask, const T *u, Epi<T> ep,
                                                     T *__restrict__ work) {
    __shared__ T su[2][CNPD];     // staged node planes k and k+1 (free dofs; Dirichlet dofs read 0)
    __shared__ T eb[24 * EBS];    // the layer's element results: eb[(a*3 + i)*EBS + element]
    // EPI_PCHEBY0: a ring of two coarse node planes under the column (9x9 coarse nodes; the column's origin is
    // even, so fine node 2I lies on coarse node I)
    constexpr int CCX = UST_TX3 / 2 + 1, CCY = UST_TY3 / 2 + 1, NCP = CCX * CCY * 3;
    static_assert(NCP <= UST_TILE, "one coarse value per thread");
    __shared__ T sc[MODE == EPI_PCHEBY0 ? 2 * NCP : 1];
    const int ntxy = g.tnx * g.tny;
    const int tile = blockIdx.x;
    const int txy = tile % ntxy, chunk = tile / ntxy;
    const int x0 = (txy % g.tnx) * UST_TX3, y0 = (txy / g.tnx) * UST_TY3;
    const int z0 = chunk * g.tz, z1 = min(z0 + g.tz, g.nz);
    const int tid = threadIdx.x, lx = tid % UST_TX3, ly = tid / UST_TX3;
    const bool active = x0 + lx < g.nx && y0 + ly < g.ny;
    const unsigned pstride = (unsigned)g.nnx * (unsigned)g.nny * 3u;  // dofs of one global node plane
    const int cnx = g.nx / 2 + 1, cny = g.ny / 2 + 1, cnz = g.nz / 2 + 1;  // EPI_PCHEBY0: the coarse grid
    // the plane dofs this thread stages and finishes: t = tid + 256 r -> node n = t / 3 (i = n % 17, j = n / 17),
    // component c = t % 3; pofs = the in-plane part of the global dof index, -1 outside the plane or the grid
    int pofs[CNR];
    bool ring[CNR];
#pragma unroll
    for (int r = 0; r < CNR; r++) {
        const int t = tid + UST_TILE * r, n = t / 3, c = t % 3;
        const int gx = x0 + n % CPX, gy = y0 + n / CPX;
        const bool ok = t < CNPD && gx < g.nnx && gy < g.nny;
        pofs[r] = ok ? (gx + g.nnx * gy) * 3 + c : -1;
        ring[r] = on_tile_face(gx, UST_TX3, g.tnx) || on_tile_face(gy, UST_TY3, g.tny);
    }
    // global loads of plane gz (issued ahead of their use)
    auto load_raw = [&](int gz, T *x, unsigned char *m) {
#pragma unroll
        for (int r = 0; r < CNR; r++) {
            const unsigned dof = pofs[r] >= 0 ? (unsigned)pofs[r] + (unsigned)gz * pstride : 0u;
            m[r] = pofs[r] >= 0 ? mask[dof] : (unsigned char)1;
            x[r] = stage_load<MODE, T>(ep, u, dof);
        }
    };
    // coarse plane c under the column: thread t < NCP loads one value (0 outside the coarse grid)
    auto coarse_load = [&](int c) -> T {
        const int n = tid / 3, cc = tid % 3;
        const int I = x0 / 2 + n % CCX, J = y0 / 2 + n / CCX;
        const bool ok = tid < NCP && I < cnx && J < cny && c < cnz;
        const T v = ep.xc[ok ? ((unsigned)I + (unsigned)cnx * ((unsigned)J + (unsigned)cny * c)) * 3u + cc : 0u];
        return ok ? v : (T)0;
    };
    // staged value: u on free dofs, plus P xc from the coarse ring for EPI_PCHEBY0
    auto staged = [&](int gz, const T *x, const unsigned char *m, T *v) {
#pragma unroll
        for (int r = 0; r < CNR; r++) {
            T val = m[r] ? (T)0 : x[r];
            if (MODE == EPI_PCHEBY0 && !m[r]) {
                const int t = tid + UST_TILE * r, n = t / 3, c = t % 3;
                val += prolong_value<T>(x0 + n % CPX, y0 + n / CPX, gz, [&](int I, int J, int K) {
                    return sc[(K & 1) * NCP + ((I - x0 / 2) + CCX * (J - y0 / 2)) * 3 + c];
                });
            }
            v[r] = val;
        }
    };
    // finish plane dof r of global plane gz with value v: the epilogue for a node of this workgroup only, the
    // workspace slot (local plane lp) for a node it shares
    T *slot = work + (size_t)tile * CNP * (g.tz + 1) * 3;
    auto finish = [&](int gz, int lp, bool zshared, const T *planeval, const EpiOps<T> *ops, const T *v) {
#pragma unroll
        for (int r = 0; r < CNR; r++) {
            if (pofs[r] < 0) continue;
            const int t = tid + UST_TILE * r;
            if (ring[r] || zshared) {
                slot[(size_t)lp * CNPD + t] = v[r];
            } else {
                EpiOps<T> o = ops[r];
                if (MODE == EPI_PCHEBY0 && !o.m) o.x = planeval[t];  // x + P xc, as staged
                epilogue_store<MODE, T>(ep, o, (unsigned)pofs[r] + (unsigned)gz * pstride, v[r]);
            }
        }
    };
    auto load_ops = [&](int gz, bool zshared, EpiOps<T> *ops) {
#pragma unroll
        for (int r = 0; r < CNR; r++) {
            const bool own = pofs[r] >= 0 && !ring[r] && !zshared;
            ops[r] = epilogue_load<MODE, T>(ep, mask, own ? (unsigned)pofs[r] + (unsigned)gz * pstride : 0u);
        }
    };
    const size_t wbase = ((size_t)txy * g.nz) * 27 * UST_TILE + tid;
    auto load_w = [&](int k, WT *w) {
#pragma unroll
        for (int q = 0; q < 27; q++) w[q] = W[wbase + ((size_t)k * 27 + q) * UST_TILE];
    };
    // when the next layer's node plane and weights are loaded (UST_COL_PF_*): wpre at the start of this layer,
    // wlate after its element apply; held: in registers (xr, mr, wr) until the next layer uses them
    constexpr int pf = col_pf(sizeof(T), MODE);
    constexpr bool wpre = pf == 1, wlate = pf == 2, held = pf != 0;
    // ---- prologue: the coarse ring (planes z0/2 and z0/2 + 1), plane z0, and the loads for layer z0 ----
    if (MODE == EPI_PCHEBY0) {
        const int c0 = z0 >> 1;
        const T v0 = coarse_load(c0), v1 = coarse_load(c0 + 1);
        if (tid < NCP) {
            sc[(c0 & 1) * NCP + tid] = v0;
            sc[((c0 + 1) & 1) * NCP + tid] = v1;
        }
        __syncthreads();
    }
    T xr[CNR], carry[CNR];
    unsigned char mr[CNR];
    {
        T v[CNR];
        load_raw(z0, xr, mr);
        staged(z0, xr, mr, v);
#pragma unroll
        for (int r = 0; r < CNR; r++) {
            const int t = tid + UST_TILE * r;
            if (t < CNPD) su[0][t] = v[r];
            carry[r] = (T)0;  // plane z0 receives nothing from below within this chunk
        }
    }
    if (held) load_raw(z0 + 1, xr, mr);
    WT wr[27];
    if (held) load_w(z0, wr);
    // EPI_PCHEBY0: staging an even plane p writes coarse plane p/2 + 1 into the ring (the odd plane p + 1 needs
    // it); its value is loaded one layer ahead
    T cv = (T)0;
    if (MODE == EPI_PCHEBY0 && wpre && ((z0 + 1) & 1) == 0) cv = coarse_load(((z0 + 1) >> 1) + 1);
    const T s[3] = {(T)(2.0 / g.hx), (T)(2.0 / g.hy), (T)(2.0 / g.hz)};
    for (int k = z0; k < z1; k++) {
        const int cur = (k - z0) & 1, nxt = cur ^ 1;
        const bool zsh = k == z0 && z0 > 0;  // the chunk's first plane is shared with the chunk below
        // this layer's operands, loaded during the previous one
        T x1[CNR];
        unsigned char m1[CNR];
        if (held) {
#pragma unroll
            for (int r = 0; r < CNR; r++) {
                x1[r] = xr[r];
                m1[r] = mr[r];
            }
        } else {
            load_raw(k + 1, x1, m1);
        }
        T wq[27];
        if (held) {
#pragma unroll
            for (int q = 0; q < 27; q++) wq[q] = wload<T, WT>(wr, q);
        }
        const T cvk = (MODE == EPI_PCHEBY0 && !wpre && ((k + 1) & 1) == 0) ? coarse_load(((k + 1) >> 1) + 1) : cv;
        // the next layer's loads
        if (wpre && k + 2 <= z1) load_raw(k + 2, xr, mr);
        if (wpre && k + 1 < z1) load_w(k + 1, wr);
        if (MODE == EPI_PCHEBY0 && wpre && ((k + 2) & 1) == 0 && k + 1 < z1) cv = coarse_load(((k + 2) >> 1) + 1);
        EpiOps<T> ops[CNR];
        if (wpre) load_ops(k, zsh, ops);
        __syncthreads();  // the previous layer is done with su[nxt], eb and the coarse ring slot written below
        {
            T v[CNR];
            staged(k + 1, x1, m1, v);
#pragma unroll
            for (int r = 0; r < CNR; r++) {
                const int t = tid + UST_TILE * r;
                if (t < CNPD) su[nxt][t] = v[r];
            }
            if (MODE == EPI_PCHEBY0 && ((k + 1) & 1) == 0 && tid < NCP)
                sc[((((k + 1) >> 1) + 1) & 1) * NCP + tid] = cvk;
        }
        __syncthreads();
        {
            if (!held) {
#pragma unroll
                for (int q = 0; q < 27; q++) wq[q] = wload<T, WT>(W, wbase + ((size_t)k * 27 + q) * UST_TILE);
            }
            T ye[8][3];
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int i = 0; i < 3; i++) ye[a][i] = (T)0;
            const T *sb = &su[cur][(lx + CPX * ly) * 3], *st = &su[nxt][(lx + CPX * ly) * 3];
            auto ue = [&](int a, int i) -> T {
                return (a >> 2 ? st : sb)[((a & 1) + CPX * ((a >> 1) & 1)) * 3 + i];
            };
            element_apply3<T>(ue, wq, s, ye);
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int i = 0; i < 3; i++) eb[(a * 3 + i) * EBS + tid] = active ? ye[a][i] : (T)0;
        }
        if (wlate && k + 2 <= z1) load_raw(k + 2, xr, mr);
        if (wlate && k + 1 < z1) load_w(k + 1, wr);
        if (!wpre) load_ops(k, zsh, ops);
        __syncthreads();
        // plane k: the carried contributions of layer k-1 plus those of layer k below the plane; the
        // contributions of layer k above plane k+1 become the new carry.  Fixed order: elements (i-1, j-1),
        // (i, j-1), (i-1, j), (i, j).
        T v[CNR];
#pragma unroll
        for (int r = 0; r < CNR; r++) {
            const int t = tid + UST_TILE * r, n = t / 3, c = t % 3, i = n % CPX, j = n / CPX;
            T lo = (T)0, hi = (T)0;
            if (t < CNPD) {
#pragma unroll
                for (int dy = 1; dy >= 0; dy--)
#pragma unroll
                    for (int dx = 1; dx >= 0; dx--) {
                        const int ex = i - dx, ey = j - dy;
                        if (ex >= 0 && ex < UST_TX3 && ey >= 0 && ey < UST_TY3) {
                            const int a = dx + 2 * dy, e = ex + UST_TX3 * ey;
                            lo += eb[(a * 3 + c) * EBS + e];
                            hi += eb[((a + 4) * 3 + c) * EBS + e];
                        }
                    }
            }
            v[r] = carry[r] + lo;
            carry[r] = hi;
        }
        finish(k, k - z0, zsh, su[cur], ops, v);
    }
    // the chunk's last plane: shared with the chunk above unless it is the top of the grid
    {
        const bool zsh = z1 < g.nz;
        EpiOps<T> ops[CNR];
        load_ops(z1, zsh, ops);
        finish(z1, z1 - z0, zsh, su[(z1 - z0) & 1], ops, carry);
    }
}

// ------------------------------------------------------------------------------------------
// 3D tile kernel for grids too small for the column kernel to fill the GPU (the coarse multigrid levels):
// one 256-thread workgroup per 16x4x4 tile, one z layer of 64 elements per wavefront, the node block staged
// and reduced in LDS (FP64 accumulation, two wave-parity phases fix the summation order), tile-face nodes
// through the workspace.
// ------------------------------------------------------------------------------------------
template <typename T, typename WT, int MODE>
__global__ __launch_bounds__(256, 2) void k_applyK3_tiled(GridDev g, const WT *__restrict__ W,
                                                       const unsigned char *__restrict__ mask, const T *u,
                                                       Epi<T> ep, T *__restrict__ work) {
    constexpr int TX = UST_TTX3, TY = UST_TTY3, TZ = UST_TTZ3;
    constexpr int NX = TX + 1, NY = TY + 1, NZ = TZ + 1, NN = NX * NY * NZ;
    __shared__ T su[NN * 3];
    __shared__ double sy[NN * 3];  // FP64 accumulation in LDS for both precisions (ds_add_f64)
    // EPI_PCHEBY0: the coarse nodes under the tile (tile origins are even, so fine node 2I lies on coarse I)
    constexpr int CX = TX / 2 + 1, CY = TY / 2 + 1, CZ = TZ / 2 + 1, NC = CX * CY * CZ * 3;
    static_assert(NC <= UST_TILE, "one coarse value per thread");
    __shared__ T sc[MODE == EPI_PCHEBY0 ? NC : 1];
    const int tile = blockIdx.x;
    const int tx0 = (tile % g.tnx) * TX, ty0 = ((tile / g.tnx) % g.tny) * TY, tz0 = (tile / (g.tnx * g.tny)) * TZ;
    T cv = (T)0;
    if (MODE == EPI_PCHEBY0) {
        const int t = threadIdx.x, c = t % 3, n = t / 3;
        const int I = tx0 / 2 + n % CX, J = ty0 / 2 + (n / CX) % CY, K = tz0 / 2 + n / (CX * CY);
        const int cnx = g.nx / 2 + 1, cny = g.ny / 2 + 1, cnz = g.nz / 2 + 1;
        const bool ok = t < NC && I < cnx && J < cny && K < cnz;
        cv = ep.xc[ok ? ((unsigned)I + (unsigned)cnx * ((unsigned)J + (unsigned)cny * K)) * 3u + c : 0u];
        cv = ok ? cv : (T)0;
    }
    // Loads are issued without branches, staged velocities first, then the element weights (which do not
    // depend on them), so the LDS stores below wait only for the staging loads and the weights arrive
    // while the tile is being staged.  Out-of-tile slots read dof 0 and are discarded; partial tiles are
    // padded in W, so every load is in bounds.
    constexpr int NST = (NN * 3 + UST_TILE - 1) / UST_TILE;
    T stv[NST];
    bool stok[NST];
    unsigned char stm[NST];
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        const int c = t % 3, n = t / 3;
        const int gx = tx0 + n % NX, gy = ty0 + (n / NX) % NY, gz = tz0 + n / (NX * NY);
        stok[r] = t < NN * 3 && gx < g.nnx && gy < g.nny && gz < g.nnz;
        const unsigned dof = stok[r] ? ((unsigned)gx + (unsigned)g.nnx * ((unsigned)gy + (unsigned)g.nny * gz)) * 3u + c : 0u;
        stm[r] = mask[dof];
        stv[r] = stage_load<MODE, T>(ep, u, dof);
    }
    // FP32: the weights are loaded here, ahead of the barrier; FP64 loads them after it, which keeps the
    // register count within two waves per SIMD
    constexpr bool early = sizeof(T) == 4;
    T wq[27];
    if (early) {
#pragma unroll
        for (int q = 0; q < 27; q++) wq[q] = wload<T, WT>(W, ((size_t)tile * 27 + q) * UST_TILE + threadIdx.x);
    }
    if (MODE == EPI_PCHEBY0) {
        if (threadIdx.x < NC) sc[threadIdx.x] = cv;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NST; r++) {
            const int t = threadIdx.x + UST_TILE * r;
            const int c = t % 3, n = t / 3;
            const int ix = n % NX, iy = (n / NX) % NY, iz = n / (NX * NY);
            if (stok[r] && !stm[r])
                stv[r] += prolong_value<T>(ix, iy, iz, [&](int I, int J, int K) {
                    return sc[(I + CX * (J + CY * K)) * 3 + c];
                });
        }
    }
#pragma unroll
    for (int r = 0; r < NST; r++) stv[r] = (stok[r] && !stm[r]) ? stv[r] : (T)0;
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        if (t < NN * 3) {
            su[t] = stv[r];
            sy[t] = 0.0;
        }
    }
    __syncthreads();
    const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
    const bool active = tx0 + lx < g.nx && ty0 + ly < g.ny && tz0 + lz < g.nz;
    T ye[8][3];
    if (active) {
        if (!early) {
#pragma unroll
            for (int q = 0; q < 27; q++) wq[q] = wload<T, WT>(W, ((size_t)tile * 27 + q) * UST_TILE + threadIdx.x);
        }
        T ue[8][3];
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const int ln = (lx + (a & 1)) + NX * ((ly + ((a >> 1) & 1)) + NY * (lz + (a >> 2)));
#pragma unroll
            for (int i = 0; i < 3; i++) {
                ue[a][i] = su[ln * 3 + i];
                ye[a][i] = (T)0;
            }
        }
        const T s[3] = {(T)(2.0 / g.hx), (T)(2.0 / g.hy), (T)(2.0 / g.hz)};
        element_apply3<T>([&](int a, int i) -> T { return ue[a][i]; }, wq, s, ye);
    }
    // Wave w holds the element layer lz = w, whose nodes are layers w and w+1: even waves never share a
    // node, nor do odd waves, and within one LDS add the 64 lanes hit 64 distinct nodes.  Adding in two
    // wave-parity phases therefore fixes the order in which every node receives its contributions.
    for (int phase = 0; phase < 2; phase++) {
        if (active && (lz & 1) == phase) {
#pragma unroll
            for (int a = 0; a < 8; a++) {
                const int ln = (lx + (a & 1)) + NX * ((ly + ((a >> 1) & 1)) + NY * (lz + (a >> 2)));
#pragma unroll
                for (int i = 0; i < 3; i++) atomicAdd(&sy[ln * 3 + i], (double)ye[a][i]);
            }
        }
        __syncthreads();
    }
    // writeout: every load of the thread (mask, epilogue operands) first, then the stores.  Dof indices are
    // 32-bit (the host rejects larger grids), which keeps the per-thread index registers down.
    T *slot = work + (size_t)tile * NN * 3;
    int kind[NST];  // 0 none, 1 tile-face node (workspace slot), 2 node of this tile only (written here)
    unsigned wdof[NST];
    EpiOps<T> ops[NST];
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        const int c = t % 3, n = t / 3;
        const int gx = tx0 + n % NX, gy = ty0 + (n / NX) % NY, gz = tz0 + n / (NX * NY);
        const bool in = t < NN * 3 && gx < g.nnx && gy < g.nny && gz < g.nnz;
        const bool face = on_tile_face(gx, TX, g.tnx) || on_tile_face(gy, TY, g.tny) || on_tile_face(gz, TZ, g.tnz);
        kind[r] = in ? (face ? 1 : 2) : 0;
        wdof[r] = kind[r] == 2 ? ((unsigned)gx + (unsigned)g.nnx * ((unsigned)gy + (unsigned)g.nny * gz)) * 3u + c : 0u;
        ops[r] = epilogue_load<MODE, T>(ep, mask, wdof[r]);
    }
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        if (MODE == EPI_PCHEBY0 && kind[r] == 2 && !ops[r].m) ops[r].x = su[t];  // x + P xc, as staged
        if (kind[r] == 1) slot[t] = (T)sy[t];
        else if (kind[r] == 2) epilogue_store<MODE, T>(ep, ops[r], wdof[r], (T)sy[t]);
    }
}

// ------------------------------------------------------------------------------------------
// 3D halo-tile kernel (FP32, the small coarse multigrid levels): a 256-thread workgroup owns an 8x4x4 block of
// NODES and computes every element touching them -- its 8x4x4 elements and a one-element halo on the low side
// of each axis, 9x5x5 slots, one per thread -- so every owned node's sum is complete inside the workgroup: no
// workspace and no gather launch.  The halo elements are computed twice (by two workgroups), which costs less
// on the latency-bound small grids than the gather pass it replaces.  Neighbouring workgroups stage the nodes
// a workgroup writes, so the Chebyshev modes must run out of place (ep.xs != ep.y).  Each node sums its 8
// element contributions in a fixed order: deterministic.
// ------------------------------------------------------------------------------------------
constexpr int HOX = 8, HOY = 4, HOZ = 4;                    // owned nodes per axis
constexpr int HEX = HOX + 1, HEY = HOY + 1, HEZ = HOZ + 1;  // element slots per axis
constexpr int HNX = HOX + 2, HNY = HOY + 2, HNZ = HOZ + 2;  // staged nodes per axis
constexpr int HNE = HEX * HEY * HEZ, HNN = HNX * HNY * HNZ;  // 225 element slots, 360 staged nodes
constexpr int HND = HOX * HOY * HOZ * 3;                    // 384 owned dofs
static_assert(HNE <= UST_TILE && HND <= 2 * UST_TILE, "one element and at most two owned dofs per thread");

// Grids up to this many elements take the halo-tile kernel; larger tile grids keep tiles + gather.  Measured on
// the 128x128x64 solve (same box, graph-replayed): the halo passes on the 32x32x16 and 16x16x8 levels take it
// from 0.431-0.434 s to 0.403-0.407 s; on the 64x64x32 level as well, 0.415-0.417 s (the duplicated halo
// elements cost more there than the gather).  A 16x4x4-node block (two element slots per thread) was slower.
#ifndef UST_HALO_MAX_NEL
#define UST_HALO_MAX_NEL 65536
#endif

__host__ __device__ inline int halo_tiles(const GridDev &g) {
    return ((g.nnx + HOX - 1) / HOX) * ((g.nny + HOY - 1) / HOY) * ((g.nnz + HOZ - 1) / HOZ);
}

template <typename T, typename WT, int MODE>
__global__ __launch_bounds__(256) void k_applyK3_halo(GridDev g, const WT *__restrict__ W,
                                                      const unsigned char *__restrict__ mask, const T *u, Epi<T> ep) {
    __shared__ T su[HNN * 3];   // staged node block, origin (nx0 - 1, ny0 - 1, nz0 - 1)
    __shared__ T eb[24 * HNE];  // element results: eb[(a*3 + i) * HNE + slot] (HNE odd: rows shift banks)
    const int hnx = (g.nnx + HOX - 1) / HOX, hny = (g.nny + HOY - 1) / HOY;
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int nx0 = (tile % hnx) * HOX, ny0 = ((tile / hnx) % hny) * HOY, nz0 = (tile / (hnx * hny)) * HOZ;
    const int cnx = g.nx / 2 + 1, cny = g.ny / 2 + 1;  // EPI_PCHEBY0: the coarse grid
    // staging: value t of the block is node n = t / 3, component c = t % 3; nodes outside the grid stage 0
    constexpr int NST = (HNN * 3 + UST_TILE - 1) / UST_TILE;
    T stv[NST];
    unsigned char stm[NST];
    bool stok[NST];
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = tid + UST_TILE * r, c = t % 3, n = t / 3;
        const int gx = nx0 - 1 + n % HNX, gy = ny0 - 1 + (n / HNX) % HNY, gz = nz0 - 1 + n / (HNX * HNY);
        stok[r] = t < HNN * 3 && gx >= 0 && gy >= 0 && gz >= 0 && gx < g.nnx && gy < g.nny && gz < g.nnz;
        const unsigned dof = stok[r] ? ((unsigned)gx + (unsigned)g.nnx * ((unsigned)gy + (unsigned)g.nny * gz)) * 3u + c
                                     : 0u;
        stm[r] = mask[dof];
        stv[r] = stage_load<MODE, T>(ep, u, dof);
    }
    // the thread's element slot and its weights, loaded while the block is staged
    const int lx = tid % HEX, ly = (tid / HEX) % HEY, lz = tid / (HEX * HEY);
    const int ex = nx0 - 1 + lx, ey = ny0 - 1 + ly, ez = nz0 - 1 + lz;
    const bool valid = tid < HNE && ex >= 0 && ey >= 0 && ez >= 0 && ex < g.nx && ey < g.ny && ez < g.nz;
    const int e = valid ? ex + g.nx * (ey + g.ny * ez) : 0;
    T wq[27];
#pragma unroll
    for (int q = 0; q < 27; q++) wq[q] = wload<T, WT>(W, widx(g, e, q));
    if (MODE == EPI_PCHEBY0) {  // x + P xc on the free dofs, the coarse values read from global memory
#pragma unroll
        for (int r = 0; r < NST; r++) {
            const int t = tid + UST_TILE * r, c = t % 3, n = t / 3;
            const int gx = nx0 - 1 + n % HNX, gy = ny0 - 1 + (n / HNX) % HNY, gz = nz0 - 1 + n / (HNX * HNY);
            if (stok[r] && !stm[r])
                stv[r] += prolong_value<T>(gx, gy, gz, [&](int I, int J, int K) {
                    return ep.xc[((unsigned)I + (unsigned)cnx * ((unsigned)J + (unsigned)cny * K)) * 3u + c];
                });
        }
    }
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = tid + UST_TILE * r;
        if (t < HNN * 3) su[t] = (stok[r] && !stm[r]) ? stv[r] : (T)0;
    }
    __syncthreads();
    if (tid < HNE) {
        T ye[8][3];
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) ye[a][i] = (T)0;
        const T s[3] = {(T)(2.0 / g.hx), (T)(2.0 / g.hy), (T)(2.0 / g.hz)};
        const T *sb = &su[(lx + HNX * (ly + HNY * lz)) * 3];
        element_apply3<T>([&](int a, int i) -> T {
            return sb[((a & 1) + HNX * (((a >> 1) & 1) + HNY * (a >> 2))) * 3 + i];
        }, wq, s, ye);
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int i = 0; i < 3; i++) eb[(a * 3 + i) * HNE + tid] = valid ? ye[a][i] : (T)0;
    }
    __syncthreads();
    // the owned dofs d = tid, tid + 256 (< 384): node d / 3 of the block, component d % 3; each sums its 8
    // elements in a fixed order, then the epilogue
    EpiOps<T> ops[2];
    unsigned dofs[2];
    bool own[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int d = tid + UST_TILE * r, n = d / 3, c = d % 3;
        const int gx = nx0 + n % HOX, gy = ny0 + (n / HOX) % HOY, gz = nz0 + n / (HOX * HOY);
        own[r] = d < HND && gx < g.nnx && gy < g.nny && gz < g.nnz;
        dofs[r] = own[r] ? ((unsigned)gx + (unsigned)g.nnx * ((unsigned)gy + (unsigned)g.nny * gz)) * 3u + c : 0u;
        ops[r] = epilogue_load<MODE, T>(ep, mask, dofs[r]);
    }
#pragma unroll
    for (int r = 0; r < 2; r++) {
        if (!own[r]) continue;
        const int d = tid + UST_TILE * r, n = d / 3, c = d % 3;
        const int ox = n % HOX, oy = (n / HOX) % HOY, oz = n / (HOX * HOY);
        T v = (T)0;
#pragma unroll
        for (int dz = 1; dz >= 0; dz--)
#pragma unroll
            for (int dy = 1; dy >= 0; dy--)
#pragma unroll
                for (int dx = 1; dx >= 0; dx--)  // element slot (ox + 1 - dx, ...) has this node as vertex (dx, dy, dz)
                    v += eb[((dx + 2 * dy + 4 * dz) * 3 + c) * HNE + (ox + 1 - dx) + HEX * ((oy + 1 - dy) + HEY * (oz + 1 - dz))];
        EpiOps<T> o = ops[r];
        if (MODE == EPI_PCHEBY0 && !o.m) o.x = su[((ox + 1) + HNX * ((oy + 1) + HNY * (oz + 1))) * 3 + c];  // as staged
        epilogue_store<MODE, T>(ep, o, dofs[r], v);
    }
}

template <typename T, typename WT, int MODE>
__global__ __launch_bounds__(256) void k_applyK2_tiled(GridDev g, const WT *__restrict__ W,
                                                       const unsigned char *__restrict__ mask, const T *u,
                                                       Epi<T> ep, T *__restrict__ work) {
    constexpr int TX = UST_TX2, TY = UST_TY2;
    constexpr int NX = TX + 1, NY = TY + 1, NN = NX * NY;
    __shared__ T su[NN * 2];
    __shared__ double sy[NN * 2];
    const int tile = blockIdx.x;
    const int tx0 = (tile % g.tnx) * TX, ty0 = (tile / g.tnx) * TY;
    constexpr int NST = (NN * 2 + UST_TILE - 1) / UST_TILE;
    T stv[NST];
    bool stok[NST];
    unsigned char stm[NST];
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        const int c = t % 2, n = t / 2;
        const int gx = tx0 + n % NX, gy = ty0 + n / NX;
        stok[r] = t < NN * 2 && gx < g.nnx && gy < g.nny;
        const unsigned dof = stok[r] ? ((unsigned)gx + (unsigned)g.nnx * gy) * 2u + c : 0u;
        stm[r] = mask[dof];
        stv[r] = stage_load<MODE, T>(ep, u, dof);
    }
    T wq[9];
#pragma unroll
    for (int q = 0; q < 9; q++) wq[q] = wload<T, WT>(W, ((size_t)tile * 9 + q) * UST_TILE + threadIdx.x);
#pragma unroll
    for (int r = 0; r < NST; r++) stv[r] = (stok[r] && !stm[r]) ? stv[r] : (T)0;
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        if (t < NN * 2) {
            su[t] = stv[r];
            sy[t] = 0.0;
        }
    }
    __syncthreads();
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const bool active = tx0 + lx < g.nx && ty0 + ly < g.ny;
    T ye[4][2];
    if (active) {
        T ue[4][2];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int ln = (lx + (a & 1)) + NX * (ly + (a >> 1));
            ue[a][0] = su[ln * 2];
            ue[a][1] = su[ln * 2 + 1];
            ye[a][0] = ye[a][1] = (T)0;
        }
        const T s[3] = {(T)(2.0 / g.hx), (T)(2.0 / g.hy), (T)1};
        element_apply2<T>(ue, wq, s, ye);
    }
    // wave w holds element rows 4w..4w+3 (node rows 4w..4w+4): same two-phase ordering as in 3D
    const int wave = threadIdx.x / 64;
    for (int phase = 0; phase < 2; phase++) {
        if (active && (wave & 1) == phase) {
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int ln = (lx + (a & 1)) + NX * (ly + (a >> 1));
                atomicAdd(&sy[ln * 2], (double)ye[a][0]);
                atomicAdd(&sy[ln * 2 + 1], (double)ye[a][1]);
            }
        }
        __syncthreads();
    }
    T *slot = work + (size_t)tile * NN * 2;
    int kind[NST];
    unsigned wdof[NST];
    EpiOps<T> ops[NST];
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        const int c = t % 2, n = t / 2;
        const int gx = tx0 + n % NX, gy = ty0 + n / NX;
        const bool in = t < NN * 2 && gx < g.nnx && gy < g.nny;
        const bool face = on_tile_face(gx, TX, g.tnx) || on_tile_face(gy, TY, g.tny);
        kind[r] = in ? (face ? 1 : 2) : 0;
        wdof[r] = kind[r] == 2 ? ((unsigned)gx + (unsigned)g.nnx * gy) * 2u + c : 0u;
        ops[r] = epilogue_load<MODE, T>(ep, mask, wdof[r]);
    }
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int t = threadIdx.x + UST_TILE * r;
        if (kind[r] == 1) slot[t] = (T)sy[t];
        else if (kind[r] == 2) epilogue_store<MODE, T>(ep, ops[r], wdof[r], (T)sy[t]);
    }
}

// tiles holding node coordinate gc along one axis as (tile, local index) pairs: 1 or 2 of them
__device__ __forceinline__ int node_tiles(int gc, int T, int tn, int *tt, int *ll) {
    int t = gc / T;
    if (t >= tn) t = tn - 1;
    tt[0] = t;
    ll[0] = gc - t * T;
    if (ll[0] == 0 && t > 0) {
        tt[1] = t - 1;
        ll[1] = T;
        return 2;
    }
    return 1;
}

// Phase 2 visits only the nodes on tile faces, enumerated as three disjoint regular sets so that every
// thread has work: Z = the z-face planes (all x, y), Y = the y-face planes off the z-faces, X = the
// x-face lines off both.  An axis with tn tiles has tn - 1 face coordinates T, 2T, ..., (tn-1)T.
__device__ __forceinline__ int face_coord(int m, int T) { return T * (m + 1); }

// k-th coordinate that is not a tile face: 0, 1..T-1, T+1..2T-1, ..., (tn-1)T+1 .. end
__device__ __forceinline__ int nonface_coord(int k, int T, int tn) {
    if (k == 0) return 0;
    const int kk = k - 1, blk = kk / (T - 1);
    if (blk < tn - 1) return 1 + blk * T + kk % (T - 1);
    return 1 + (tn - 1) * T + (kk - (tn - 1) * (T - 1));
}

template <int DIM>
static inline unsigned face_node_count(const GridDev &g) {
    const unsigned fx = g.tnx - 1, fy = g.tny - 1, fz = DIM == 3 ? g.tnz - 1 : 0;
    return fz * g.nnx * g.nny + fy * g.nnx * (g.nnz - fz) + fx * (g.nny - fy) * (g.nnz - fz);
}

// one thread per tile-face dof: consecutive threads take the components of a node, so the slot reads of a
// wavefront are contiguous (32-bit index arithmetic: grids stay below 2^31 dofs).  Tiles are g.tx x g.ty (x g.tz)
// elements; a workgroup's slot holds its (tx+1)(ty+1)(tz+1) node block.
template <typename T, int DIM, int MODE>
__global__ __launch_bounds__(256) void k_gather_tiles(GridDev g, const unsigned char *__restrict__ mask,
                                                      const T *__restrict__ work, Epi<T> ep) {
    const int TX = g.tx, TY = g.ty, TZ = DIM == 3 ? g.tz : 1;
    const int NX = TX + 1, NY = TY + 1, NN = NX * NY * (DIM == 3 ? TZ + 1 : 1);
    const int fx = g.tnx - 1, fy = g.tny - 1, fz = DIM == 3 ? g.tnz - 1 : 0;
    const unsigned nZ = (unsigned)fz * g.nnx * g.nny;
    const unsigned nY = (unsigned)fy * g.nnx * (g.nnz - fz);
    const unsigned nX = (unsigned)fx * (g.nny - fy) * (g.nnz - fz);
    const unsigned tt = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (int)(tt % DIM);
    unsigned t = tt / DIM;
    int gx, gy, gz;
    if (t < nZ) {
        const unsigned r = t / g.nnx;
        gx = (int)(t - r * g.nnx);
        gy = (int)(r % g.nny);
        gz = face_coord((int)(r / g.nny), TZ);
    } else if ((t -= nZ) < nY) {
        const unsigned r = t / g.nnx;
        gx = (int)(t - r * g.nnx);
        gy = face_coord((int)(r % fy), TY);
        gz = DIM == 3 ? nonface_coord((int)(r / fy), TZ, g.tnz) : 0;
    } else if ((t -= nY) < nX) {
        const unsigned r = t / fx, ny = g.nny - fy;
        gx = face_coord((int)(t - r * fx), TX);
        gy = nonface_coord((int)(r % ny), TY, g.tny);
        gz = DIM == 3 ? nonface_coord((int)(r / ny), TZ, g.tnz) : 0;
    } else {
        return;
    }
    const unsigned n = (unsigned)gx + (unsigned)g.nnx * ((unsigned)gy + (unsigned)g.nny * gz);
    int tx[2], lxs[2], ty[2], lys[2], tz[2] = {0, 0}, lzs[2] = {0, 0};
    const int nxs = node_tiles(gx, TX, g.tnx, tx, lxs);
    const int nys = node_tiles(gy, TY, g.tny, ty, lys);
    const int nzs = DIM == 3 ? node_tiles(gz, TZ, g.tnz, tz, lzs) : 1;
    EpiOps<T> ops = epilogue_load<MODE, T>(ep, mask, n * DIM + c);
    T acc = (T)0;
    for (int a = 0; a < nzs; a++)
        for (int b = 0; b < nys; b++)
            for (int d = 0; d < nxs; d++) {
                const unsigned tile = (unsigned)tx[d] + (unsigned)g.tnx * ((unsigned)ty[b] + (unsigned)g.tny * tz[a]);
                const unsigned ln = lxs[d] + NX * (lys[b] + NY * lzs[a]);
                acc += work[((size_t)tile * NN + ln) * DIM + c];
            }
    if (MODE == EPI_PCHEBY0 && !ops.m) {  // 3D only (host-checked): x + P xc on a free component
        const int cnx = g.nx / 2 + 1, cny = g.ny / 2 + 1;
        ops.x += prolong_value<T>(gx, gy, gz, [&](int I, int J, int K) {
            return ep.xc[((unsigned)I + (unsigned)cnx * ((unsigned)J + (unsigned)cny * K)) * 3u + c];
        });
    }
    epilogue_store<MODE, T>(ep, ops, n * DIM + c, acc);
}

static inline size_t work_len(const GridDev &g) {
    const size_t nn = g.dim == 3 ? (size_t)(g.tx + 1) * (g.ty + 1) * (g.tz + 1) : (size_t)(g.tx + 1) * (g.ty + 1);
    return (size_t)g.ntiles * nn * g.dim;
}

template <typename T, typename WT, int MODE>
static int apply_tiled(const ust_grid *gr, const WT *W, const unsigned char *mask, const T *u, Epi<T> ep, T *work,
                       void *stream) {
    GridDev g = grid_dev(gr, col_wg_target(sizeof(T), MODE));
    if (!W || !mask || !u || !ep.y || !work) return UST_EINVAL;
    if (MODE != EPI_Y && !ep.b) return UST_EINVAL;
    if (cheby_mode(MODE) && (!ep.invd || (MODE == EPI_CHEBY && !ep.d) || ep.xs != u)) return UST_EINVAL;
    if (MODE == EPI_RES0 && (!ep.invd || !ep.d)) return UST_EINVAL;
    if (MODE == EPI_PCHEBY0 && (g.dim != 3 || !ep.xc || g.nx % 2 || g.ny % 2 || g.nz % 2)) return UST_EINVAL;
    if ((size_t)g.nnx * g.nny * g.nnz * g.dim >= (1ull << 32)) return UST_EINVAL;  // 32-bit dof indices
    hipStream_t st = ust_stream(stream);
    // FP32 on a small tile grid: the halo-tile kernel, unless a Chebyshev step runs in place
    const bool halo = sizeof(T) == 4 && g.dim == 3 && !g.col && g.nel <= UST_HALO_MAX_NEL &&
                      (!cheby_mode(MODE) || ep.y != ep.xs);
    if (halo) {
        if constexpr (sizeof(T) == 4)
            hipLaunchKernelGGL((k_applyK3_halo<T, WT, MODE>), dim3(halo_tiles(g)), dim3(UST_TILE), 0, st, g, W, mask,
                               u, ep);
    } else if (g.dim == 3) {
        if (g.col)
            hipLaunchKernelGGL((k_applyK3_col<T, WT, MODE>), dim3(g.ntiles), dim3(UST_TILE), 0, st, g, W, mask, u, ep,
                               work);
        else
            hipLaunchKernelGGL((k_applyK3_tiled<T, WT, MODE>), dim3(g.ntiles), dim3(UST_TILE), 0, st, g, W, mask, u,
                               ep, work);
        const unsigned nf = face_node_count<3>(g);
        if (nf) hipLaunchKernelGGL((k_gather_tiles<T, 3, MODE>), dim3(ust_grid_size((size_t)nf * 3, 256)), dim3(256),
                                   0, st, g, mask, work, ep);
    } else {
        hipLaunchKernelGGL((k_applyK2_tiled<T, WT, MODE>), dim3(g.ntiles), dim3(UST_TILE), 0, st, g, W, mask, u, ep,
                           work);
        const unsigned nf = face_node_count<2>(g);
        if (nf) hipLaunchKernelGGL((k_gather_tiles<T, 2, MODE>), dim3(ust_grid_size((size_t)nf * 2, 256)), dim3(256),
                                   0, st, g, mask, work, ep);
    }
    UST_LAUNCH_CHECK();
    return UST_OK;
}

template <typename T, typename WT>
static int apply_plain(const ust_grid *g, const WT *W, const unsigned char *mask, const T *u, T *y, T *work,
                       void *stream) {
    Epi<T> ep = {y, nullptr, nullptr, nullptr, (T)0, (T)0, nullptr, nullptr};
    return apply_tiled<T, WT, EPI_Y>(g, W, mask, u, ep, work, stream);
}

template <typename T, typename WT>
static int apply_residual(const ust_grid *g, const WT *W, const unsigned char *mask, const T *u, const T *b, T *r,
                          T *work, void *stream) {
    Epi<T> ep = {r, b, nullptr, nullptr, (T)0, (T)0, nullptr, nullptr};
    return apply_tiled<T, WT, EPI_RES>(g, W, mask, u, ep, work, stream);
}

// xi: x before the step; x: after it (the same vector when the step runs in place)
template <typename T, typename WT>
static int apply_cheby(const ust_grid *g, const WT *W, const unsigned char *mask, const T *xi, T *x, const T *b,
                       const T *invd, T *d, T c1, T c2, int first, T *work, void *stream) {
    Epi<T> ep = {x, b, invd, d, c1, c2, nullptr, xi};
    return first ? apply_tiled<T, WT, EPI_CHEBY0>(g, W, mask, xi, ep, work, stream)
                 : apply_tiled<T, WT, EPI_CHEBY>(g, W, mask, xi, ep, work, stream);
}

template <typename T, typename WT>
static int apply_cheby_prolong(const ust_grid *g, const WT *W, const unsigned char *mask, const T *xi, T *x,
                               const T *b, const T *invd, T *d, T c2, const T *xc, T *work, void *stream) {
    Epi<T> ep = {x, b, invd, d, (T)0, c2, xc, xi};
    return apply_tiled<T, WT, EPI_PCHEBY0>(g, W, mask, xi, ep, work, stream);
}

template <typename T, typename WT>
static int apply_residual_first(const ust_grid *g, const WT *W, const unsigned char *mask, const T *b, const T *invd,
                                T *x, T *r, T c2, T *work, void *stream) {
    Epi<T> ep = {r, b, invd, x, (T)0, c2, nullptr, nullptr};
    return apply_tiled<T, WT, EPI_RES0>(g, W, mask, b, ep, work, stream);  // u unused: b stands in
}

extern "C" int ust_k_apply_K_residual_first_f32(const ust_grid *g, const float *WK, const unsigned char *mask,
                                                const float *b, const float *invd, float *x, float *r, float c2,
                                                float *work, void *stream) {
    return apply_residual_first<float, float>(g, WK, mask, b, invd, x, r, c2, work, stream);
}

// the largest workspace any pass needs (the passes' column grids differ in z-chunk length)
extern "C" size_t ust_k_apply_work_len(const ust_grid *gr) {
    const size_t a = work_len(grid_dev(gr, UST_COL_WG_TARGET)), b = work_len(grid_dev(gr, UST_COL_WG_TARGET_F32_YR));
    return a > b ? a : b;
}

extern "C" int ust_k_apply_K_cheby_prolong_f64(const ust_grid *g, const double *WK, const unsigned char *mask,
                                               double *x, const double *b, const double *invd, double *d, double c2,
                                               const double *xc, double *work, void *stream) {
    return apply_cheby_prolong<double, double>(g, WK, mask, x, x, b, invd, d, c2, xc, work, stream);
}

extern "C" int ust_k_apply_K_cheby_prolong_f32(const ust_grid *g, const float *WK, const unsigned char *mask,
                                               float *x, const float *b, const float *invd, float *d, float c2,
                                               const float *xc, float *work, void *stream) {
    return apply_cheby_prolong<float, float>(g, WK, mask, x, x, b, invd, d, c2, xc, work, stream);
}

// out of place: x before the step, xo after it (distinct vectors); on FP32 tile grids this runs the halo-tile
// kernel (one pass, no gather)
extern "C" int ust_k_apply_K_cheby_oop_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *x,
                                           float *xo, const float *b, const float *invd, float *d, float c1, float c2,
                                           int first, float *work, void *stream) {
    if (!x || x == xo) return UST_EINVAL;
    return apply_cheby<float, float>(g, WK, mask, x, xo, b, invd, d, c1, c2, first, work, stream);
}

extern "C" int ust_k_apply_K_cheby_prolong_oop_f32(const ust_grid *g, const float *WK, const unsigned char *mask,
                                                   const float *x, float *xo, const float *b, const float *invd,
                                                   float *d, float c2, const float *xc, float *work, void *stream) {
    if (!x || x == xo) return UST_EINVAL;
    return apply_cheby_prolong<float, float>(g, WK, mask, x, xo, b, invd, d, c2, xc, work, stream);
}

extern "C" int ust_k_apply_K_f64(const ust_grid *g, const double *WK, const unsigned char *mask, const double *u,
                                 double *y, double *work, void *stream) {
    return apply_plain<double, double>(g, WK, mask, u, y, work, stream);
}

extern "C" int ust_k_apply_K_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *u,
                                 float *y, float *work, void *stream) {
    return apply_plain<float, float>(g, WK, mask, u, y, work, stream);
}

extern "C" int ust_k_apply_K_residual_f64(const ust_grid *g, const double *WK, const unsigned char *mask,
                                          const double *u, const double *b, double *r, double *work, void *stream) {
    return apply_residual<double, double>(g, WK, mask, u, b, r, work, stream);
}

extern "C" int ust_k_apply_K_residual_f32(const ust_grid *g, const float *WK, const unsigned char *mask,
                                          const float *u, const float *b, float *r, float *work, void *stream) {
    return apply_residual<float, float>(g, WK, mask, u, b, r, work, stream);
}

extern "C" int ust_k_apply_K_cheby_f64(const ust_grid *g, const double *WK, const unsigned char *mask, double *x,
                                       const double *b, const double *invd, double *d, double c1, double c2,
                                       int first, double *work, void *stream) {
    return apply_cheby<double, double>(g, WK, mask, x, x, b, invd, d, c1, c2, first, work, stream);
}

extern "C" int ust_k_apply_K_cheby_f32(const ust_grid *g, const float *WK, const unsigned char *mask, float *x,
                                       const float *b, const float *invd, float *d, float c1, float c2, int first,
                                       float *work, void *stream) {
    return apply_cheby<float, float>(g, WK, mask, x, x, b, invd, d, c1, c2, first, work, stream);
}

// FP32 operator with bfloat16 K weights (the multigrid V-cycle's operator; half the weight stream)
extern "C" int ust_k_apply_K_f32_bf16w(const ust_grid *g, const unsigned short *WK, const unsigned char *mask,
                                       const float *u, float *y, float *work, void *stream) {
    return apply_plain<float, unsigned short>(g, WK, mask, u, y, work, stream);
}

extern "C" int ust_k_apply_K_residual_f32_bf16w(const ust_grid *g, const unsigned short *WK,
                                                const unsigned char *mask, const float *u, const float *b, float *r,
                                                float *work, void *stream) {
    return apply_residual<float, unsigned short>(g, WK, mask, u, b, r, work, stream);
}

extern "C" int ust_k_apply_K_cheby_f32_bf16w(const ust_grid *g, const unsigned short *WK, const unsigned char *mask,
                                             float *x, const float *b, const float *invd, float *d, float c1, float c2,
                                             int first, float *work, void *stream) {
    return apply_cheby<float, unsigned short>(g, WK, mask, x, x, b, invd, d, c1, c2, first, work, stream);
}

extern "C" int ust_k_apply_K_cheby_prolong_f32_bf16w(const ust_grid *g, const unsigned short *WK,
                                                     const unsigned char *mask, float *x, const float *b,
                                                     const float *invd, float *d, float c2, const float *xc,
                                                     float *work, void *stream) {
    return apply_cheby_prolong<float, unsigned short>(g, WK, mask, x, x, b, invd, d, c2, xc, work, stream);
}

extern "C" size_t ust_k_weights_len(const ust_grid *gr) {
    GridDev g = grid_dev(gr);
    if (g.dim == 3 && g.col) return (size_t)g.tnx * g.tny * g.nz * 27 * UST_TILE;
    if (g.dim == 3) return (size_t)g.ntiles * 27 * UST_TILE;
    return (size_t)g.ntiles * 9 * UST_TILE;
}

extern "C" size_t ust_k_weights_index(const ust_grid *gr, int e, int q) {
    GridDev g = grid_dev(gr);
    return widx(g, e, q);
}

// ------------------------------------------------------------------------------------------
// K u without a caller workspace, for callers that cannot allocate (code under HIP graph capture, serving
// loops): the per-tile node partials go to a file-scope device array.  Each GPU the library runs on holds its
// own copy until the process ends; its device address is looked up once per device and kept.  Every slot the
// gather pass reads is written by the same call's tile pass first, so nothing needs clearing between calls;
// calls that overlap on one GPU share the array and race.
// ------------------------------------------------------------------------------------------
#define UST_NOALLOC_WORK (size_t(1) << 22)  // doubles: 32 MiB
#define UST_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[UST_NOALLOC_WORK];
static double *g_scratch_dev[UST_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_address(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        ust_set_hip_error(e, __FILE__, __LINE__);
        return UST_EDEVICE;
    }
    if (dev < 0 || dev >= UST_NOALLOC_MAX_DEVICES) {
        ust_set_error_text("ust_k_apply_K_f64_noalloc: device index beyond the library's table");
        return UST_EDEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ust_set_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, __FILE__, __LINE__);
            return UST_EDEVICE;
        }
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return UST_OK;
}

extern "C" size_t ust_k_apply_noalloc_max_work(void) { return UST_NOALLOC_WORK; }

extern "C" int ust_k_apply_K_f64_noalloc(const ust_grid *g, const double *WK, const unsigned char *mask,
                                         const double *u, double *y, void *stream) {
    if (!g) return UST_EINVAL;
    const size_t need = work_len(grid_dev(g, col_wg_target(sizeof(double), EPI_Y)));
    if (need > UST_NOALLOC_WORK) {
        char msg[256];
        snprintf(msg, sizeof(msg),
                 "ust_k_apply_K_f64_noalloc: the grid needs %zu workspace doubles, more than the %zu of the "
                 "library's static workspace (use ust_k_apply_K_f64 with a workspace)",
                 need, (size_t)UST_NOALLOC_WORK);
        ust_set_error_text(msg);
        return UST_EINVAL;
    }
    double *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != UST_OK) return rc;
    return apply_plain<double, double>(g, WK, mask, u, y, ws, stream);
}
