// The hot operator of the BSSCR solve on gfx950: y = K u, matrix-free, from the 3^dim quadrature
// weights of each element (the MatMult of the A11 block, auglag-driver-DGTGD.c:191-403, that every
// inner Krylov iteration and every multigrid smoothing step performs, PETScMGSolver.c:641-716).
//
// One 256-thread workgroup owns a tile of elements (8x8x4 in 3D, 16x16 in 2D), one element per
// thread.  The tile's nodal velocities are staged in LDS, each thread applies its element with sum
// factorisation (gradients at the Lagrange points from edge differences, stress, and the transposed
// interpolation back to the vertices, ~0.8 kflop per 3D element), element results are reduced in
// LDS, and the tile writes its whole node block to its own slot of a workspace.  A second, node-
// ordered pass sums the slots covering each dof (one for interior nodes, two to eight on tile faces)
// and writes y, so y needs no clearing and no global atomics are used.
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
template <typename T, int I, int K>
__device__ __forceinline__ void grad3(const T u[8][3], T sk, T gp[3][3]) {
    T D[2][2];
#pragma unroll
    for (int b1 = 0; b1 < 2; b1++)
#pragma unroll
        for (int b2 = 0; b2 < 2; b2++)
            D[b1][b2] = (T)0.5 * sk * (u[node3(K, 1, b1, b2)][I] - u[node3(K, 0, b1, b2)][I]);
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

template <typename T, int I>
__device__ __forceinline__ void diag3(const T u[8][3], const T W[27], const T s[3], T y[8][3]) {
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

template <typename T, int I, int K>
__device__ __forceinline__ void pair3(const T u[8][3], const T W[27], const T s[3], T y[8][3]) {
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

template <typename T>
__device__ __forceinline__ void element_apply3(const T u[8][3], const T W[27], const T s[3], T y[8][3]) {
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

template <typename T>
__global__ __launch_bounds__(256) void k_applyK3_tiled(GridDev g, const T *__restrict__ W,
                                                       const unsigned char *__restrict__ mask,
                                                       const T *__restrict__ u, T *__restrict__ y,
                                                       T *__restrict__ work) {
    constexpr int TX = UST_TX3, TY = UST_TY3, TZ = UST_TZ3;
    constexpr int NX = TX + 1, NY = TY + 1, NZ = TZ + 1, NN = NX * NY * NZ;
    __shared__ T su[NN * 3];
    __shared__ T sy[NN * 3];
    const int tile = blockIdx.x;
    const int tx0 = (tile % g.tnx) * TX, ty0 = ((tile / g.tnx) % g.tny) * TY, tz0 = (tile / (g.tnx * g.tny)) * TZ;
    for (int t = threadIdx.x; t < NN * 3; t += blockDim.x) {
        const int c = t % 3, n = t / 3;
        const int gx = tx0 + n % NX, gy = ty0 + (n / NX) % NY, gz = tz0 + n / (NX * NY);
        T v = (T)0;
        if (gx < g.nnx && gy < g.nny && gz < g.nnz) {
            const size_t dof = ((size_t)gx + (size_t)g.nnx * ((size_t)gy + (size_t)g.nny * gz)) * 3 + c;
            if (!mask[dof]) v = u[dof];
        }
        su[t] = v;
        sy[t] = (T)0;
    }
    __syncthreads();
    const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
    const bool active = tx0 + lx < g.nx && ty0 + ly < g.ny && tz0 + lz < g.nz;
    T ye[8][3];
    if (active) {
        T ue[8][3], wq[27];
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const int ln = (lx + (a & 1)) + NX * ((ly + ((a >> 1) & 1)) + NY * (lz + (a >> 2)));
#pragma unroll
            for (int i = 0; i < 3; i++) {
                ue[a][i] = su[ln * 3 + i];
                ye[a][i] = (T)0;
            }
        }
#pragma unroll
        for (int q = 0; q < 27; q++) wq[q] = W[((size_t)tile * 27 + q) * UST_TILE + threadIdx.x];
        const T s[3] = {(T)(2.0 / g.hx), (T)(2.0 / g.hy), (T)(2.0 / g.hz)};
        element_apply3<T>(ue, wq, s, ye);
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
                for (int i = 0; i < 3; i++) atomicAdd(&sy[ln * 3 + i], ye[a][i]);
            }
        }
        __syncthreads();
    }
    T *slot = work + (size_t)tile * NN * 3;
    for (int t = threadIdx.x; t < NN * 3; t += blockDim.x) {
        const int c = t % 3, n = t / 3;
        const int gx = tx0 + n % NX, gy = ty0 + (n / NX) % NY, gz = tz0 + n / (NX * NY);
        if (gx >= g.nnx || gy >= g.nny || gz >= g.nnz) continue;
        if (on_tile_face(gx, TX, g.tnx) || on_tile_face(gy, TY, g.tny) || on_tile_face(gz, TZ, g.tnz)) {
            slot[t] = sy[t];
        } else {
            const size_t dof = ((size_t)gx + (size_t)g.nnx * ((size_t)gy + (size_t)g.nny * gz)) * 3 + c;
            y[dof] = mask[dof] ? (T)0 : sy[t];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_applyK2_tiled(GridDev g, const T *__restrict__ W,
                                                       const unsigned char *__restrict__ mask,
                                                       const T *__restrict__ u, T *__restrict__ y,
                                                       T *__restrict__ work) {
    constexpr int TX = UST_TX2, TY = UST_TY2;
    constexpr int NX = TX + 1, NY = TY + 1, NN = NX * NY;
    __shared__ T su[NN * 2];
    __shared__ T sy[NN * 2];
    const int tile = blockIdx.x;
    const int tx0 = (tile % g.tnx) * TX, ty0 = (tile / g.tnx) * TY;
    for (int t = threadIdx.x; t < NN * 2; t += blockDim.x) {
        const int c = t % 2, n = t / 2;
        const int gx = tx0 + n % NX, gy = ty0 + n / NX;
        T v = (T)0;
        if (gx < g.nnx && gy < g.nny) {
            const size_t dof = ((size_t)gx + (size_t)g.nnx * gy) * 2 + c;
            if (!mask[dof]) v = u[dof];
        }
        su[t] = v;
        sy[t] = (T)0;
    }
    __syncthreads();
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const bool active = tx0 + lx < g.nx && ty0 + ly < g.ny;
    T ye[4][2];
    if (active) {
        T ue[4][2], wq[9];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int ln = (lx + (a & 1)) + NX * (ly + (a >> 1));
            ue[a][0] = su[ln * 2];
            ue[a][1] = su[ln * 2 + 1];
            ye[a][0] = ye[a][1] = (T)0;
        }
#pragma unroll
        for (int q = 0; q < 9; q++) wq[q] = W[((size_t)tile * 9 + q) * UST_TILE + threadIdx.x];
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
                atomicAdd(&sy[ln * 2], ye[a][0]);
                atomicAdd(&sy[ln * 2 + 1], ye[a][1]);
            }
        }
        __syncthreads();
    }
    T *slot = work + (size_t)tile * NN * 2;
    for (int t = threadIdx.x; t < NN * 2; t += blockDim.x) {
        const int c = t % 2, n = t / 2;
        const int gx = tx0 + n % NX, gy = ty0 + n / NX;
        if (gx >= g.nnx || gy >= g.nny) continue;
        if (on_tile_face(gx, TX, g.tnx) || on_tile_face(gy, TY, g.tny)) {
            slot[t] = sy[t];
        } else {
            const size_t dof = ((size_t)gx + (size_t)g.nnx * gy) * 2 + c;
            y[dof] = mask[dof] ? (T)0 : sy[t];
        }
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

template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_gather_tiles(GridDev g, const unsigned char *__restrict__ mask,
                                                      const T *__restrict__ work, T *__restrict__ y) {
    const size_t ndof = (size_t)g.nnx * g.nny * g.nnz * DIM;
    const size_t dof = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (dof >= ndof) return;
    const int c = (int)(dof % DIM);
    const size_t n = dof / DIM;
    const int gx = (int)(n % g.nnx), gy = (int)((n / g.nnx) % g.nny), gz = (int)(n / ((size_t)g.nnx * g.nny));
    constexpr int TX = DIM == 3 ? UST_TX3 : UST_TX2, TY = DIM == 3 ? UST_TY3 : UST_TY2, TZ = DIM == 3 ? UST_TZ3 : 1;
    constexpr int NX = TX + 1, NY = TY + 1, NN = NX * NY * (DIM == 3 ? TZ + 1 : 1);
    int tx[2], lxs[2], ty[2], lys[2], tz[2] = {0, 0}, lzs[2] = {0, 0};
    const int nxs = node_tiles(gx, TX, g.tnx, tx, lxs);
    const int nys = node_tiles(gy, TY, g.tny, ty, lys);
    const int nzs = DIM == 3 ? node_tiles(gz, TZ, g.tnz, tz, lzs) : 1;
    if (nxs * nys * nzs == 1) return;  // written by its only tile
    if (mask[dof]) {
        y[dof] = (T)0;
        return;
    }
    T acc = (T)0;
    for (int a = 0; a < nzs; a++)
        for (int b = 0; b < nys; b++)
            for (int d = 0; d < nxs; d++) {
                const size_t tile = (size_t)tx[d] + (size_t)g.tnx * ((size_t)ty[b] + (size_t)g.tny * tz[a]);
                const int ln = lxs[d] + NX * (lys[b] + NY * lzs[a]);
                acc += work[(tile * NN + ln) * DIM + c];
            }
    y[dof] = acc;
}

static inline size_t work_len(const GridDev &g) {
    const int nn = g.dim == 3 ? (UST_TX3 + 1) * (UST_TY3 + 1) * (UST_TZ3 + 1) : (UST_TX2 + 1) * (UST_TY2 + 1);
    return (size_t)g.ntiles * nn * g.dim;
}

template <typename T>
static int apply_tiled(const ust_grid *gr, const T *W, const unsigned char *mask, const T *u, T *y, T *work,
                       void *stream) {
    GridDev g = grid_dev(gr);
    if (!W || !mask || !u || !y || !work) return UST_EINVAL;
    const size_t ndof = (size_t)g.nnx * g.nny * g.nnz * g.dim;
    hipStream_t st = ust_stream(stream);
    if (g.dim == 3) {
        hipLaunchKernelGGL(k_applyK3_tiled<T>, dim3(g.ntiles), dim3(UST_TILE), 0, st, g, W, mask, u, y, work);
        hipLaunchKernelGGL((k_gather_tiles<T, 3>), dim3(ust_grid_size(ndof, 256)), dim3(256), 0, st, g, mask, work, y);
    } else {
        hipLaunchKernelGGL(k_applyK2_tiled<T>, dim3(g.ntiles), dim3(UST_TILE), 0, st, g, W, mask, u, y, work);
        hipLaunchKernelGGL((k_gather_tiles<T, 2>), dim3(ust_grid_size(ndof, 256)), dim3(256), 0, st, g, mask, work, y);
    }
    UST_LAUNCH_CHECK();
    return UST_OK;
}

extern "C" size_t ust_k_apply_work_len(const ust_grid *gr) { return work_len(grid_dev(gr)); }

extern "C" int ust_k_apply_K_f64(const ust_grid *g, const double *WK, const unsigned char *mask, const double *u,
                                 double *y, double *work, void *stream) {
    return apply_tiled<double>(g, WK, mask, u, y, work, stream);
}

extern "C" int ust_k_apply_K_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *u,
                                 float *y, float *work, void *stream) {
    return apply_tiled<float>(g, WK, mask, u, y, work, stream);
}

extern "C" size_t ust_k_weights_len(const ust_grid *gr) {
    GridDev g = grid_dev(gr);
    return (size_t)g.ntiles * (g.dim == 3 ? 27 : 9) * UST_TILE;
}

extern "C" size_t ust_k_weights_index(const ust_grid *gr, int e, int q) {
    GridDev g = grid_dev(gr);
    return widx(g, e, q);
}
