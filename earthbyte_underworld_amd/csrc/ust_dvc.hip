// Discrete Voronoi cell (DVC) integration weights on gfx950.
//
// DVCWeights (PICellerator/Weights/src/DVCWeights.c): every element carries a fixed sub-grid of nx*ny(*nz)
// cells over its local box [-1,1]^dim; the particles grow their Voronoi regions over it from their seed cells,
// one particle after another, round after round, until no particle claims a cell.  A particle claims an unowned
// cell on its boundary, and takes an owned one when it is strictly nearer the cell centre than the owner
// (_ClaimCells :261-317 with _DistanceTest :604-610); its next boundary is the set of neighbours of the cells it
// has just claimed that it does not own (_UpdateBchain :395-446).  Then each particle moves to the centroid of
// its cells and weighs their volume (_GetCentroids :203-231, _Calculate3D/2D :763-925).
//
// Within one particle's turn every boundary cell is decided on its own (owner and distances only), and the next
// boundary is a set, so the reference's lists can be held as sets: one bit per sub-cell.  Only the particle
// order is sequential.  One 64-lane wave owns one element and a sub-cell set is 64-bit words, word j covering
// sub-cells 64 j .. 64 j + 63.  In a turn, lane L decides sub-cell 64 j + L of every boundary word j that holds
// cells, one ballot per word, which leaves the claimed set one word per lane; the next boundary is that set
// dilated by whole-word shifts across lanes (+-1, +-nx, +-nx*ny sub-cells, masked at the sub-grid's x and y
// faces), less the cells the particle owns.  The boundary sets of all particles of the element, the owner grid
// and the particle coordinates stay in LDS.
#include "ust_common.h"

namespace {

constexpr int DVC_LANES = 64;

struct DvcShape {
    int n[3];     // sub-grid resolution (nz = 1 in 2D)
    int ncell;    // n[0] n[1] n[2]
    int words;    // 64-bit words of a sub-cell set
    int max_np;   // particles per element the LDS layout holds
    float inv[2]; // 1 / n[0], 1 / n[1]
};

// floor(c / n) for 0 <= c < 4096 and 1 <= n <= 64: (c + 1/2) / n is at least 1/128 away from an integer and the
// float product errs by less than 4096 * 2^-22 < 1/1000, so the truncation is exact
__device__ __forceinline__ int small_div(int c, float inv) { return (int)(((float)c + 0.5f) * inv); }

// the (i, j, k) of sub-cell c (x fastest) without integer divisions
struct CellWalk {
    int i, j, k;
    __device__ __forceinline__ CellWalk(const DvcShape &s, int c) {
        const int t = small_div(c, s.inv[0]);
        i = c - t * s.n[0];
        k = small_div(t, s.inv[1]);
        j = t - k * s.n[1];
    }
};

// (x_o + x_p - 2 x_c).(x_o - x_p) in the reference's evaluation order, without contraction into fused
// multiply-adds: > 0 when particle p is strictly nearer the centre c than the owner o
template <int DIM>
__device__ __forceinline__ double dvc_distance_test(const double *c, const double *po, const double *pp) {
#pragma clang fp contract(off)
    double s = (po[0] + pp[0] - c[0] - c[0]) * (po[0] - pp[0]);
    s = s + (po[1] + pp[1] - c[1] - c[1]) * (po[1] - pp[1]);
    if (DIM == 3) s = s + (po[2] + pp[2] - c[2] - c[2]) * (po[2] - pp[2]);
    return s;
}

// 64-bit value of lane `src` (0 when src is outside the wave)
__device__ __forceinline__ unsigned long long lane_word(unsigned long long x, int src) {
    const int lo = __shfl((int)(unsigned)x, src & 63), hi = __shfl((int)(unsigned)(x >> 32), src & 63);
    const unsigned long long v = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
    return (src >= 0 && src < 64) ? v : 0ull;
}

// a sub-cell set held one word per lane (lane t: sub-cells 64 t .. 64 t + 63), shifted by s sub-cells:
// bit c of up(x, s) is bit c - s of x, bit c of down(x, s) is bit c + s of x
__device__ __forceinline__ unsigned long long set_up(unsigned long long x, int s, int lane) {
    const int q = s >> 6, r = s & 63;
    const unsigned long long a = lane_word(x, lane - q);
    return r ? (a << r) | (lane_word(x, lane - q - 1) >> (64 - r)) : a;
}
__device__ __forceinline__ unsigned long long set_down(unsigned long long x, int s, int lane) {
    const int q = s >> 6, r = s & 63;
    const unsigned long long a = lane_word(x, lane + q);
    return r ? (a >> r) | (lane_word(x, lane + q + 1) << (64 - r)) : a;
}

// flag bits: 1 element without particles, 2 more particles than the LDS layout holds, 4 round limit reached.
// OT: the owner grid's type, signed char when at most 127 particles share an element (less LDS, more waves)
template <int DIM, typename OT>
__global__ __launch_bounds__(DVC_LANES) void k_dvc_weights(DvcShape s, const int *cell_off, int nel, double *xi0,
                                                           double *xi1, double *xi2, double *w, int set_centroids,
                                                           int *flag) {
    extern __shared__ __align__(16) unsigned char dvc_lds[];
    const int e = blockIdx.x;
    const int lane = threadIdx.x;
    if (e >= nel) return;
    const int off = cell_off[e];
    const int np = cell_off[e + 1] - off;
    if (np <= 0 || np > s.max_np) {
        if (lane == 0) atomicOr(flag, np <= 0 ? 1 : 2);
        return;
    }
    const int W = s.words, ncell = s.ncell;
    // LDS: centres per axis | particle coordinates (np x DIM) | boundary sets (np x W) | owner
    double *axc = (double *)dvc_lds;                                  // n[0] + n[1] + n[2]
    double *pc = axc + (s.n[0] + s.n[1] + s.n[2]);                    // max_np * DIM
    unsigned long long *bset = (unsigned long long *)(pc + (size_t)s.max_np * DIM);  // max_np * W
    OT *owner = (OT *)(bset + (size_t)s.max_np * W);                  // ncell

    const double d[3] = {2.0 / s.n[0], 2.0 / s.n[1], 2.0 / s.n[2]};
    if (lane < DIM) {  // centres accumulated as the grid constructor does: -1 - d/2, then + d per cell
        const int n = lane == 0 ? s.n[0] : (lane == 1 ? s.n[1] : s.n[2]);
        const double dl = lane == 0 ? d[0] : (lane == 1 ? d[1] : d[2]);
        double *a = axc + (lane == 0 ? 0 : (lane == 1 ? s.n[0] : s.n[0] + s.n[1]));
        double x = -1.0 - dl / 2.0;
        for (int i = 0; i < n; i++) {
            x = x + dl;
            a[i] = x;
        }
    }
    for (int p = lane; p < np; p += DVC_LANES) {
        pc[p * DIM + 0] = xi0[off + p];
        pc[p * DIM + 1] = xi1[off + p];
        if (DIM == 3) pc[p * DIM + 2] = xi2[off + p];
    }
    for (int c = lane; c < ncell; c += DVC_LANES) owner[c] = -1;
    __syncthreads();

    const double *cxs = axc, *cys = axc + s.n[0], *czs = axc + s.n[0] + s.n[1];
    const int nx = s.n[0], nxy = s.n[0] * s.n[1];

    // per lane t < W, the sub-cells of word t that have a neighbour in -x, +x, -y, +y, and that exist
    unsigned long long m_im = 0ull, m_ip = 0ull, m_jm = 0ull, m_jp = 0ull, m_val = 0ull;
    if (lane < W) {
        CellWalk cw(s, 64 * lane);
        for (int b = 0; b < 64 && 64 * lane + b < ncell; b++) {
            const unsigned long long bit = 1ull << b;
            m_val |= bit;
            if (cw.i > 0) m_im |= bit;
            if (cw.i < nx - 1) m_ip |= bit;
            if (cw.j > 0) m_jm |= bit;
            if (cw.j < s.n[1] - 1) m_jp |= bit;
            if (++cw.i == nx) {
                cw.i = 0;
                if (++cw.j == s.n[1]) {
                    cw.j = 0;
                    cw.k++;
                }
            }
        }
    }

    // the new boundary of particle p from the set it has just claimed (lane t < W holds word t in `cw_set`;
    // owner final): the face neighbours of the claimed cells, by whole-word shifts, less the cells p owns
    auto update = [&](int p, unsigned long long cw_set) {
        unsigned long long nb = (set_up(cw_set, 1, lane) & m_im) | (set_down(cw_set, 1, lane) & m_ip) |
                                (set_up(cw_set, nx, lane) & m_jm) | (set_down(cw_set, nx, lane) & m_jp);
        if (DIM == 3) nb |= set_up(cw_set, nxy, lane) | set_down(cw_set, nxy, lane);
        nb &= m_val;
        unsigned long long live = __ballot(nb != 0ull);
        unsigned long long mine = 0ull;  // lane t < W ends with word t of the new boundary
        while (live) {
            const int j = __builtin_ctzll(live);
            live &= live - 1ull;
            const unsigned long long nj = lane_word(nb, j);
            const bool b = ((nj >> lane) & 1ull) && owner[64 * j + lane] != p;
            const unsigned long long m = __ballot(b);
            if (lane == j) mine = m;
        }
        if (lane < W) bset[(size_t)p * W + lane] = mine;
    };

    // seeds, each followed at once by its boundary (DVCWeights.c:676-696)
    for (int p = 0; p < np; p++) {
        int idx[3] = {0, 0, 0};  // every lane computes the seed (the same value)
#pragma unroll
        for (int a = 0; a < DIM; a++) {
            const double t = (pc[p * DIM + a] - (-1.0)) / d[a];
            // C truncation toward zero; k == n moves to n - 1; outside [-1, 1] (and NaN): the nearest cell
            idx[a] = (t >= 0.0) ? (t < (double)s.n[a] ? (int)t : s.n[a] - 1) : 0;
        }
        const int seed = idx[0] + idx[1] * nx + (DIM == 3 ? idx[2] * nxy : 0);
        if (lane == 0) owner[seed] = (OT)p;
        __syncthreads();
        update(p, lane == (seed >> 6) ? 1ull << (seed & 63) : 0ull);
        __syncthreads();
    }

    // rounds over the particles in order until nothing is claimed (DVCWeights.c:698-710).  Every claim changes
    // an owner and an owner only changes to a strictly nearer particle, so the rounds are bounded; the limit
    // below keeps the loop finite whatever the rounding does.
    const long long max_rounds = (long long)ncell * np + 2;
    long long rounds = 0;
    int total = 1;
    while (total != 0) {
        if (++rounds > max_rounds) {
            if (lane == 0) atomicOr(flag, 4);
            break;
        }
        total = 0;
        for (int p = 0; p < np; p++) {
            const unsigned long long *bp = bset + (size_t)p * W;
            const double *pp = pc + p * DIM;
            int claimed = 0;
            // the words of p's boundary that hold cells, one ballot; only those are visited
            const unsigned long long bmine = lane < W ? bp[lane] : 0ull;
            unsigned long long live = __ballot(bmine != 0ull);
            unsigned long long cmine = 0ull;  // lane t < W ends with word t of the claimed set
            while (live) {
                const int j = __builtin_ctzll(live);
                live &= live - 1ull;
                const unsigned long long bw = bp[j];
                const int c = 64 * j + lane;
                bool win = false;
                if ((bw >> lane) & 1ull) {
                    const int o = owner[c];
                    if (o == -1) {
                        win = true;
                    } else if (o != p) {
                        const CellWalk cw(s, c);
                        const double cc[3] = {cxs[cw.i], cys[cw.j], DIM == 3 ? czs[cw.k] : 0.0};
                        win = dvc_distance_test<DIM>(cc, pc + o * DIM, pp) > 0.0;
                    }
                    if (win) owner[c] = (OT)p;
                }
                const unsigned long long m = __ballot(win);
                if (lane == j) cmine = m;
                claimed += __popcll(m);
            }
            __syncthreads();
            update(p, cmine);
            __syncthreads();
            total += claimed;
        }
    }

    // centroids summed in sub-cell order, weight = count * cell volume (DVCWeights.c:203-231)
    const double vol = DIM == 3 ? d[0] * d[1] * d[2] : d[0] * d[1];
    for (int p = lane; p < np; p += DVC_LANES) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        int count = 0;
        int c = 0;
        for (int k = 0; k < (DIM == 3 ? s.n[2] : 1); k++)
            for (int j = 0; j < s.n[1]; j++)
                for (int i = 0; i < nx; i++, c++)
                    if (owner[c] == p) {
                        sx += cxs[i];
                        sy += cys[j];
                        if (DIM == 3) sz += czs[k];
                        count++;
                    }
        w[off + p] = (double)count * vol;
        if (set_centroids) {
            xi0[off + p] = count ? sx / count : 0.0;
            xi1[off + p] = count ? sy / count : 0.0;
            if (DIM == 3) xi2[off + p] = count ? sz / count : 0.0;
        }
    }
}

}  // namespace

extern "C" {

size_t ust_k_dvc_lds_bytes(int dim, const int *res, int max_np) {
    const int nz = dim == 3 ? res[2] : 1;
    const size_t ncell = (size_t)res[0] * res[1] * nz;
    const size_t words = (ncell + 63) / 64;
    return (size_t)(res[0] + res[1] + nz) * 8 + (size_t)max_np * dim * 8 + (size_t)max_np * words * 8 +
           ncell * (max_np <= 127 ? sizeof(signed char) : sizeof(short));
}

int ust_k_dvc_weights(int dim, const int *res, int nel, const int *cell_off, int max_np, double *xi0, double *xi1,
                      double *xi2, double *w, int set_centroids, int *flag, void *stream) {
    if ((dim != 2 && dim != 3) || !res || nel < 0 || !cell_off || !xi0 || !xi1 || !w || !flag || max_np < 1 ||
        max_np > 32767)
        return UST_EINVAL;
    if (dim == 3 && !xi2) return UST_EINVAL;
    DvcShape s;
    s.n[0] = res[0];
    s.n[1] = res[1];
    s.n[2] = dim == 3 ? res[2] : 1;
    for (int a = 0; a < 3; a++)
        if (s.n[a] < 1 || s.n[a] > 64) return UST_EINVAL;
    s.ncell = s.n[0] * s.n[1] * s.n[2];
    if (s.ncell > 4096) return UST_EINVAL;
    s.words = (s.ncell + 63) / 64;
    s.max_np = max_np;
    s.inv[0] = 1.0f / (float)s.n[0];
    s.inv[1] = 1.0f / (float)s.n[1];
    const size_t lds = ust_k_dvc_lds_bytes(dim, res, max_np);
    if (lds > 65536) return UST_EINVAL;
    if (nel == 0) return UST_OK;
    hipStream_t st = ust_stream(stream);
    const bool narrow = max_np <= 127;
    if (dim == 3 && narrow)
        hipLaunchKernelGGL((k_dvc_weights<3, signed char>), dim3(nel), dim3(DVC_LANES), lds, st, s, cell_off, nel,
                           xi0, xi1, xi2, w, set_centroids, flag);
    else if (dim == 3)
        hipLaunchKernelGGL((k_dvc_weights<3, short>), dim3(nel), dim3(DVC_LANES), lds, st, s, cell_off, nel, xi0,
                           xi1, xi2, w, set_centroids, flag);
    else if (narrow)
        hipLaunchKernelGGL((k_dvc_weights<2, signed char>), dim3(nel), dim3(DVC_LANES), lds, st, s, cell_off, nel,
                           xi0, xi1, xi2, w, set_centroids, flag);
    else
        hipLaunchKernelGGL((k_dvc_weights<2, short>), dim3(nel), dim3(DVC_LANES), lds, st, s, cell_off, nel, xi0,
                           xi1, xi2, w, set_centroids, flag);
    UST_LAUNCH_CHECK();
    return UST_OK;
}

}  // extern "C"
