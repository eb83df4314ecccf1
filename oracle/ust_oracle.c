/*
 * ust_oracle.c -- FP64 CPU restatement of Underworld 1.x's PIC Stokes hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and the
 * CPU-baseline leg of the benchmark may load this library; the product never does.
 *
 * Every routine states the reference file:line it restates.  Loop orders follow the
 * reference so that round-off matches it as closely as a re-implementation can.
 */
#include "ust_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------
 * Park-Miller MINSTD with Schrage's method: StgDomain/Swarm/src/Random.c:40-73.
 * The state is returned raw (not un-masked); x = min + (max-min)*state/(2^31-1).
 * ---------------------------------------------------------------------------------------- */
static long g_rng_state;
static const long RNG_RANDMAX = 2147483647L, RNG_A = 16807L, RNG_Q = 127773L, RNG_R = 2836L;
static const unsigned long RNG_MASK = 123459876UL;

void orc_rng_seed(long seed) { g_rng_state = seed ^ (long)RNG_MASK; }           /* Random.c:56-59 */
long orc_rng_state(void) { return g_rng_state; }
double orc_rng_random(void) {                                                     /* Random.c:62-69 */
    const long tmp = g_rng_state / RNG_Q;
    g_rng_state = RNG_A * (g_rng_state - tmp * RNG_Q) - RNG_R * tmp;
    if (g_rng_state < 0) g_rng_state += RNG_RANDMAX;
    return (double)g_rng_state;
}
double orc_rng_random_minmax(double mn, double mx) {                             /* Random.c:71-73 */
    return (mx - mn) * orc_rng_random() / (double)RNG_RANDMAX + mn;
}

int orc_nnodes(const ORC_Mesh *m) {
    int n = 1;
    for (int d = 0; d < m->dim; d++) n *= m->nel[d] + 1;
    return n;
}
int orc_nelements(const ORC_Mesh *m) {
    int n = 1;
    for (int d = 0; d < m->dim; d++) n *= m->nel[d];
    return n;
}

/* CartesianGenerator vertex coordinates: min + i*(max-min)/res (CartesianGenerator.c:976-1050). */
void orc_vertex(const ORC_Mesh *m, int node, double *x) {
    int idx[3] = {0, 0, 0}, r = node;
    for (int d = 0; d < m->dim; d++) {
        idx[d] = r % (m->nel[d] + 1);
        r /= (m->nel[d] + 1);
    }
    for (int d = 0; d < 3; d++) {
        if (d < m->dim)
            x[d] = m->xmin[d] + (double)idx[d] * (m->xmax[d] - m->xmin[d]) / (double)m->nel[d];
        else
            x[d] = 0.0;
    }
}

/* element -> vertex incidence, local node a = ix + 2 iy + 4 iz (CartesianGenerator.c:1005-1035). */
void orc_element_nodes(const ORC_Mesh *m, int e, int *nodes) {
    int idx[3] = {0, 0, 0}, r = e;
    for (int d = 0; d < m->dim; d++) {
        idx[d] = r % m->nel[d];
        r /= m->nel[d];
    }
    int nx = m->nel[0] + 1, ny = m->nel[1] + 1;
    int nen = 1 << m->dim;
    for (int a = 0; a < nen; a++) {
        int i = idx[0] + (a & 1), j = idx[1] + ((a >> 1) & 1), k = (m->dim == 3) ? idx[2] + ((a >> 2) & 1) : 0;
        nodes[a] = i + nx * (j + ny * k);
    }
}

/* Q1 shape functions: BilinearElementType.c:212-222, TrilinearElementType.c:249-265. */
void orc_shape_functions(int dim, const double *xi, double *N) {
    if (dim == 2) {
        double x = xi[0], y = xi[1];
        N[0] = 0.25 * (1.0 - x) * (1.0 - y);
        N[1] = 0.25 * (1.0 + x) * (1.0 - y);
        N[2] = 0.25 * (1.0 - x) * (1.0 + y);
        N[3] = 0.25 * (1.0 + x) * (1.0 + y);
    } else {
        double x = xi[0], y = xi[1], z = xi[2];
        N[0] = 0.125 * (1.0 - x) * (1.0 - y) * (1.0 - z);
        N[2] = 0.125 * (1.0 - x) * (1.0 + y) * (1.0 - z);
        N[3] = 0.125 * (1.0 + x) * (1.0 + y) * (1.0 - z);
        N[1] = 0.125 * (1.0 + x) * (1.0 - y) * (1.0 - z);
        N[4] = 0.125 * (1.0 - x) * (1.0 - y) * (1.0 + z);
        N[6] = 0.125 * (1.0 - x) * (1.0 + y) * (1.0 + z);
        N[7] = 0.125 * (1.0 + x) * (1.0 + y) * (1.0 + z);
        N[5] = 0.125 * (1.0 + x) * (1.0 - y) * (1.0 + z);
    }
}

/* Q1 local derivatives: BilinearElementType.c:225-243, TrilinearElementType.c:268-304. */
void orc_shape_local_derivs(int dim, const double *xi, double GNi[3][8]) {
    if (dim == 2) {
        double x = xi[0], y = xi[1];
        GNi[0][0] = -0.25 * (1.0 - y);
        GNi[0][1] = 0.25 * (1.0 - y);
        GNi[0][3] = 0.25 * (1.0 + y);
        GNi[0][2] = -0.25 * (1.0 + y);
        GNi[1][0] = -0.25 * (1.0 - x);
        GNi[1][1] = -0.25 * (1.0 + x);
        GNi[1][3] = 0.25 * (1.0 + x);
        GNi[1][2] = 0.25 * (1.0 - x);
    } else {
        double x = xi[0], y = xi[1], z = xi[2];
        GNi[0][0] = -0.125 * (1.0 - y) * (1.0 - z);
        GNi[0][2] = -0.125 * (1.0 + y) * (1.0 - z);
        GNi[0][3] = 0.125 * (1.0 + y) * (1.0 - z);
        GNi[0][1] = 0.125 * (1.0 - y) * (1.0 - z);
        GNi[0][4] = -0.125 * (1.0 - y) * (1.0 + z);
        GNi[0][6] = -0.125 * (1.0 + y) * (1.0 + z);
        GNi[0][7] = 0.125 * (1.0 + y) * (1.0 + z);
        GNi[0][5] = 0.125 * (1.0 - y) * (1.0 + z);

        GNi[1][0] = -0.125 * (1.0 - x) * (1.0 - z);
        GNi[1][2] = 0.125 * (1.0 - x) * (1.0 - z);
        GNi[1][3] = 0.125 * (1.0 + x) * (1.0 - z);
        GNi[1][1] = -0.125 * (1.0 + x) * (1.0 - z);
        GNi[1][4] = -0.125 * (1.0 - x) * (1.0 + z);
        GNi[1][6] = 0.125 * (1.0 - x) * (1.0 + z);
        GNi[1][7] = 0.125 * (1.0 + x) * (1.0 + z);
        GNi[1][5] = -0.125 * (1.0 + x) * (1.0 + z);

        GNi[2][0] = -0.125 * (1.0 - x) * (1.0 - y);
        GNi[2][2] = -0.125 * (1.0 - x) * (1.0 + y);
        GNi[2][3] = -0.125 * (1.0 + x) * (1.0 + y);
        GNi[2][1] = -0.125 * (1.0 + x) * (1.0 - y);
        GNi[2][4] = 0.125 * (1.0 - x) * (1.0 - y);
        GNi[2][6] = 0.125 * (1.0 - x) * (1.0 + y);
        GNi[2][7] = 0.125 * (1.0 + x) * (1.0 + y);
        GNi[2][5] = 0.125 * (1.0 + x) * (1.0 - y);
    }
}

/* Solve a dim x dim system by Cramer's rule (TensorArray_SolveSystem role). */
static void solve_small(int dim, double J[3][3], const double *rhs, double *x) {
    if (dim == 2) {
        double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        x[0] = (rhs[0] * J[1][1] - J[0][1] * rhs[1]) / det;
        x[1] = (J[0][0] * rhs[1] - rhs[0] * J[1][0]) / det;
    } else {
        double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        for (int c = 0; c < 3; c++) {
            double M[3][3];
            memcpy(M, J, sizeof(M));
            for (int r = 0; r < 3; r++) M[r][c] = rhs[r];
            double d = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
            x[c] = d / det;
        }
    }
}

/* Newton global->local map: ElementType.c:227-350 (xi0 = 0, tol 1e-4, 100 its, fail => xi[0] = 1.1). */
void orc_global_to_local(const ORC_Mesh *m, int e, const double *x, double *xi) {
    int dim = m->dim, nen = 1 << dim, nodes[8];
    double vx[8][3];
    orc_element_nodes(m, e, nodes);
    for (int a = 0; a < nen; a++) orc_vertex(m, nodes[a], vx[a]);
    for (int d = 0; d < dim; d++) xi[d] = 0.0;
    for (int it = 0; it < 100; it++) {
        double J[3][3] = {{0}}, rhs[3] = {0, 0, 0}, N[8], GNi[3][8], inc[3] = {0, 0, 0};
        orc_shape_functions(dim, xi, N);
        orc_shape_local_derivs(dim, xi, GNi);
        for (int a = 0; a < nen; a++)
            for (int i = 0; i < dim; i++) {
                for (int j = 0; j < dim; j++) J[i][j] += GNi[j][a] * vx[a][i];
                rhs[i] -= N[a] * vx[a][i];
            }
        for (int i = 0; i < dim; i++) rhs[i] += x[i];
        solve_small(dim, J, rhs, inc);
        double mx = 0.0;
        for (int d = 0; d < dim; d++) {
            xi[d] += inc[d];
            if (fabs(inc[d]) > mx) mx = fabs(inc[d]);
        }
        if (mx < 1e-4) return;
    }
    xi[0] = 1.1;
}

/* ElementType_ShapeFunctionsGlobalDerivs: ElementType.c:354-495 (cofactor inverse). */
void orc_shape_global_derivs(const ORC_Mesh *m, int e, const double *xi, double *detJ, double GNx[3][8]) {
    int dim = m->dim, nen = 1 << dim, nodes[8];
    double GNi[3][8], jac[3][3], cof[3][3], D, vx[3];
    orc_element_nodes(m, e, nodes);
    orc_shape_local_derivs(dim, xi, GNi);
    memset(jac, 0, sizeof(jac));
    for (int n = 0; n < nen; n++) {
        orc_vertex(m, nodes[n], vx);
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) jac[i][j] = jac[i][j] + GNi[i][n] * vx[j];
    }
    if (dim == 2) {
        D = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
        double tmp = jac[0][0];
        jac[0][0] = jac[1][1] / D;
        jac[1][1] = tmp / D;
        jac[0][1] = -jac[0][1] / D;
        jac[1][0] = -jac[1][0] / D;
    } else {
        D = jac[0][0] * (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]) -
            jac[0][1] * (jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]) +
            jac[0][2] * (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]);
        cof[0][0] = jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1];
        cof[1][0] = -(jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]);
        cof[2][0] = jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0];
        cof[0][1] = -(jac[0][1] * jac[2][2] - jac[0][2] * jac[2][1]);
        cof[1][1] = jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0];
        cof[2][1] = -(jac[0][0] * jac[2][1] - jac[0][1] * jac[2][0]);
        cof[0][2] = jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1];
        cof[1][2] = -(jac[0][0] * jac[1][2] - jac[0][2] * jac[1][0]);
        cof[2][2] = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) jac[i][j] = cof[i][j] / D;
    }
    *detJ = D;
    for (int dx = 0; dx < dim; dx++)
        for (int n = 0; n < nen; n++) {
            double v = 0.0;
            for (int dxi = 0; dxi < dim; dxi++) v = v + GNi[dxi][n] * jac[dx][dxi];
            GNx[dx][n] = v;
        }
}

int orc_find_element(const ORC_Mesh *m, const double *x) {
    int e = 0, stride = 1;
    for (int d = 0; d < m->dim; d++) {
        double h = (m->xmax[d] - m->xmin[d]) / m->nel[d];
        int i = (int)floor((x[d] - m->xmin[d]) / h);
        if (i < 0) i = 0;
        if (i >= m->nel[d]) i = m->nel[d] - 1;
        e += i * stride;
        stride *= m->nel[d];
    }
    return e;
}

/* RandomParticleLayout: cells in order, ppc per cell, 3 draws per particle even in 2D
 * (RandomParticleLayout.c:195-213) inside the cell box (Swarm_GetCellMinMaxCoords). */
/* RandomParticleLayout with a per-cell particle count (cellParticleCountTbl, RandomParticleLayout.c:195-213):
 * the cells in order, the particles of a cell in order, three draws per particle in the cell's box. */
void orc_layout_random_counts(const ORC_Mesh *m, const int *cell_off, long seed, double *coords) {
    int nel = orc_nelements(m), nodes[8];
    orc_rng_seed(seed);
    for (int e = 0; e < nel; e++) {
        double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, v0[3], v1[3];
        orc_element_nodes(m, e, nodes);
        orc_vertex(m, nodes[0], v0);
        orc_vertex(m, nodes[(1 << m->dim) - 1], v1);
        for (int d = 0; d < m->dim; d++) {
            mn[d] = v0[d];
            mx[d] = v1[d];
        }
        for (int p = cell_off[e]; p < cell_off[e + 1]; p++)
            for (int d = 0; d < 3; d++) coords[3 * (size_t)p + d] = orc_rng_random_minmax(mn[d], mx[d]);
    }
}

void orc_layout_random(const ORC_Mesh *m, int ppc, long seed, double *coords) {
    int nel = orc_nelements(m);
    int *off = (int *)malloc(sizeof(int) * (size_t)(nel + 1));
    for (int e = 0; e <= nel; e++) off[e] = e * ppc;
    orc_layout_random_counts(m, off, seed, coords);
    free(off);
}

/* CoincidentMapper (CoincidentMapper.c:150-216): xi by Newton; ConstantWeights
 * (ConstantWeights.c:180-189, cellLocalVolume = 2^dim, WeightsCalculator.c:83). */
void orc_map_particles(const ORC_Mesh *m, int np, const double *coords, const int *cell_off, double *xi,
                       double *w) {
    int nel = orc_nelements(m);
    (void)np;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nel; e++) {
        int n = cell_off[e + 1] - cell_off[e];
        double wt = (double)(1 << m->dim) / (double)n;
        for (int p = cell_off[e]; p < cell_off[e + 1]; p++) {
            double x3[3] = {0, 0, 0};
            for (int d = 0; d < m->dim; d++) x3[d] = coords[3 * (size_t)p + d];
            orc_global_to_local(m, e, x3, &xi[3 * (size_t)p]);
            if (m->dim == 2) xi[3 * (size_t)p + 2] = 0.0;
            w[p] = wt;
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Rheology chain for one particle: RheologyMaterial_RunRheologies (RheologyMaterial.c:240-270)
 * -> Rheology_Register_RunRheologies (Rheology_Register.c:61-77), registration order.
 * D stays isotropic for every rheology here, so it is carried as the scalar eta
 * (ConstitutiveMatrix_SetIsotropicViscosity / _IsotropicCorrection / _GetIsotropicViscosity,
 * ConstitutiveMatrix.c:362-392, ConstitutiveMatrixCartesian.c:449-481).
 * ---------------------------------------------------------------------------------------- */
static double second_invariant(int dim, const double *t) {                   /* TensorMath.c:430-460 */
    double inv;
    if (dim == 3)
        inv = 0.5 * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) + t[3] * t[3] + t[4] * t[4] + t[5] * t[5];
    else
        inv = 0.5 * (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
    return sqrt(inv);
}

/* strain rate (xx,yy,[zz],xy,[xz,yz]) at xi: symmetric part of the interpolated velocity gradient
 * (TensorMath.h:93-99 ordering). */
static void particle_strain_rate(const ORC_Mesh *m, int e, const double GNx[3][8], const double *vel,
                                 double *sr) {
    int dim = m->dim, nen = 1 << dim, nodes[8];
    double L[3][3] = {{0}};
    orc_element_nodes(m, e, nodes);
    for (int a = 0; a < nen; a++)
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) L[i][j] += GNx[j][a] * vel[(size_t)nodes[a] * dim + i];
    if (dim == 2) {
        sr[0] = L[0][0];
        sr[1] = L[1][1];
        sr[2] = 0.5 * (L[0][1] + L[1][0]);
    } else {
        sr[0] = L[0][0];
        sr[1] = L[1][1];
        sr[2] = L[2][2];
        sr[3] = 0.5 * (L[0][1] + L[1][0]);
        sr[4] = 0.5 * (L[0][2] + L[2][0]);
        sr[5] = 0.5 * (L[1][2] + L[2][1]);
    }
}

double orc_particle_viscosity(const ORC_Mesh *m, const ORC_Material *mat, int e, const double *xi,
                              const ORC_Fields *f, float coh_scale, long p, unsigned char *yielded) {
    int dim = m->dim, nen = 1 << dim, nodes[8];
    double eta = 0.0;
    orc_element_nodes(m, e, nodes);
    for (int r = 0; r < mat->nrh; r++) {
        const ORC_Rheology *rh = &mat->rh[r];
        switch (rh->type) {
        case ORC_RH_MATERIALVISCOSITY: /* MaterialViscosity.c:119-131 */
            eta = rh->p[0];
            break;
        case ORC_RH_ARRHENIUS: { /* Arrhenius.c:156-200 */
            double N[8], T = 0.0;
            orc_shape_functions(dim, xi, N);
            for (int a = 0; a < nen; a++) T += N[a] * f->temperature[nodes[a]];
            if (rh->p[2] > (0.0 + 1e-12)) {
                double y = 0.0, vx[3];
                for (int a = 0; a < nen; a++) {
                    orc_vertex(m, nodes[a], vx);
                    y += N[a] * vx[1];
                }
                double depth = m->xmax[1] - y;
                eta = rh->p[0] * exp((rh->p[1] + rh->p[2] * depth) / (T + rh->p[3]));
            } else {
                eta = rh->p[0] * exp((rh->p[1]) / (T + rh->p[3]));
            }
            break;
        }
        case ORC_RH_VISCOSITYLIMITER: /* ViscosityLimiter.c:106-125 */
            if (eta > rh->p[0]) eta = rh->p[0];
            if (eta < rh->p[1]) eta = rh->p[1];
            break;
        case ORC_RH_VONMISES:
        case ORC_RH_DRUCKERPRAGER: { /* YieldRheology.c:198-243 */
            int dp = (rh->type == ORC_RH_DRUCKERPRAGER);
            double minVisc = dp ? rh->p[3] : rh->p[1];
            int srSoft = (int)(dp ? rh->p[4] : rh->p[2]);
            if (!f->previous_solution_exists) break; /* :211-212 */
            double detJ, GNx[3][8], sr[6], st[6];
            orc_shape_global_derivs(m, e, xi, &detJ, GNx);
            particle_strain_rate(m, e, GNx, f->velocity, sr);
            /* ConstitutiveMatrixCartesian{2D,3D}_CalculateStress, isotropic D (:651-700) */
            int nsym = (dim == 2) ? 3 : 6;
            for (int i = 0; i < nsym; i++) st[i] = (i < dim) ? 2.0 * eta * sr[i] : eta * 2.0 * sr[i];
            double indicator = second_invariant(dim, st); /* VonMises.c:206-230 */
            /* StrainWeakening_CalcRatio (StrainWeakening.c:484-503): the stored value is clamped to >= 0 */
            int sw = f->pfw && rh->p[dp ? 7 : 4] != 0.0;
            double ratio = 0.0, pfw = 0.0;
            if (sw) {
                if (f->pfw[p] < 0.0) f->pfw[p] = 0.0;
                pfw = f->pfw[p];
                ratio = pfw / f->softening_strain;
                if (ratio > 1.0) ratio = 1.0;
            }
            /* effective cohesion, VonMises.c:194-199 / DruckerPrager.c:283-291 (the cohesion factor of the
             * swarm's noise scales the softened cohesion) */
            double cohesion = (rh->p[0] * (1.0 - ratio) + rh->p[dp ? 5 : 3] * ratio) * (double)coh_scale;
            double criterion;
            if (!dp) {
                criterion = cohesion; /* VonMises.c:182-204 */
            } else {                  /* DruckerPrager.c:236-321 */
                double friction = rh->p[1] * (1.0 - ratio) + rh->p[6] * ratio;
                double pressure = f->pressure[e], phi = atan(friction), k, c;
                if (dim == 3) {
                    double sinphi = sin(phi), oneOver = 1.0 / (sqrt(3.0) * (3.0 - sinphi));
                    k = 6.0 * sinphi * oneOver;
                    c = 6.0 * cohesion * cos(phi) * oneOver;
                } else {
                    c = cohesion * cos(phi);
                    k = sin(phi);
                }
                criterion = k * pressure + c;
                if (f->tensile) f->tensile[p] = (unsigned char)(criterion <= 0.0); /* :312 */
                if (criterion < rh->p[2]) criterion = rh->p[2];
            }
            if (sw) { /* StrainWeakening_AssignIncrement before yielding (YieldRheology.c:224-226,
                         _StrainWeakening_CalcIncrementIsotropic, StrainWeakening.c:448-482) */
                double beta = (criterion == 0.0 || indicator == 0.0) ? 0.0 : 1.0 - criterion / indicator;
                if (beta <= 0.0)
                    f->pfw_inc[p] = -f->healing_rate * pfw;
                else
                    f->pfw_inc[p] = (criterion / eta) * (beta / (1.0 - beta) - f->healing_rate * pfw);
            }
            if (yielded) *yielded = (unsigned char)(criterion < indicator);
            if (criterion < indicator) { /* VonMises.c:232-269 */
                if (srSoft) {
                    eta = 2.0 * criterion * criterion * eta / (criterion * criterion + indicator * indicator);
                } else {
                    double beta = 1.0 - criterion / indicator, corr = -eta * beta;
                    if ((eta + corr) < minVisc) corr = minVisc - eta;
                    eta += corr;
                }
            }
            if (eta < minVisc) eta = minVisc; /* YieldRheology.c:239-240 */
            break;
        }
        default:
            break;
        }
    }
    return eta;
}

void orc_strain_weakening_step(int np, double *pfw, const double *inc, double dt) {
    for (int i = 0; i < np; i++) {
        pfw[i] += dt * inc[i];
        if (pfw[i] < 0.0) pfw[i] = 0.0;
    }
}

void orc_eval_viscosity(const ORC_Mesh *m, const ORC_Material *mats, int np, const int *cell_off,
                        const double *xi, const int *mat_of, const ORC_Fields *f, double *eta,
                        unsigned char *yielded) {
    int nel = orc_nelements(m);
    (void)np;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nel; e++)
        for (int p = cell_off[e]; p < cell_off[e + 1]; p++) {
            float cs = f->coh_scale ? f->coh_scale[p] : 1.0f;
            eta[p] = orc_particle_viscosity(m, &mats[mat_of[p]], e, &xi[3 * (size_t)p], f, cs, p,
                                            yielded ? &yielded[p] : NULL);
        }
}

/* ------------------------------------------------------------------------------------------
 * _ConstitutiveMatrixCartesian_AssembleElement (ConstitutiveMatrixCartesian.c:202-412), isotropic
 * D with isDiagonal (set at :119), D~ = D*detJ*w (:335), D~B diagonal branch (:556-649),
 * B^T(D~B) accumulated into Ke[colNode*dim+i][rowNode*dim+j] (2D :354-357, 3D :388-407).
 * n particles of element e: xi (n*3), w (n), eta (n).  Ke is nd x nd row-major, accumulated.
 * ---------------------------------------------------------------------------------------- */
void orc_element_K(const ORC_Mesh *m, int e, int n, const double *xi, const double *w, const double *eta,
                   double *Ke) {
    int dim = m->dim, nen = 1 << dim, nd = nen * dim;
    for (int p = 0; p < n; p++) {
        double GNx[3][8], detJ, DB[6][3];
        orc_shape_global_derivs(m, e, &xi[3 * p], &detJ, GNx);
        double D[6];
        if (dim == 2) {
            D[0] = 2.0 * eta[p];
            D[1] = 2.0 * eta[p];
            D[2] = eta[p];
        } else {
            D[0] = D[1] = D[2] = 2.0 * eta[p];
            D[3] = D[4] = D[5] = eta[p];
        }
        int nsym = (dim == 2) ? 3 : 6;
        for (int i = 0; i < nsym; i++) D[i] *= detJ * w[p];
        for (int r = 0; r < nen; r++) {
            int rd = r * dim;
            if (dim == 2) {
                double dx = GNx[0][r], dy = GNx[1][r];
                DB[0][0] = D[0] * dx; DB[0][1] = 0.0;
                DB[1][0] = 0.0;       DB[1][1] = D[1] * dy;
                DB[2][0] = D[2] * dy; DB[2][1] = D[2] * dx;
                for (int c = 0; c < nen; c++) {
                    int cd = c * dim;
                    double bx = GNx[0][c], by = GNx[1][c];
                    Ke[(cd) * nd + rd] += bx * DB[0][0] + by * DB[2][0];
                    Ke[(cd) * nd + rd + 1] += bx * DB[0][1] + by * DB[2][1];
                    Ke[(cd + 1) * nd + rd] += by * DB[1][0] + bx * DB[2][0];
                    Ke[(cd + 1) * nd + rd + 1] += by * DB[1][1] + bx * DB[2][1];
                }
            } else {
                double dx = GNx[0][r], dy = GNx[1][r], dz = GNx[2][r];
                DB[0][0] = D[0] * dx; DB[0][1] = 0.0;       DB[0][2] = 0.0;
                DB[1][0] = 0.0;       DB[1][1] = D[1] * dy; DB[1][2] = 0.0;
                DB[2][0] = 0.0;       DB[2][1] = 0.0;       DB[2][2] = D[2] * dz;
                DB[3][0] = D[3] * dy; DB[3][1] = D[3] * dx; DB[3][2] = 0.0;
                DB[4][0] = D[4] * dz; DB[4][1] = 0.0;       DB[4][2] = D[4] * dx;
                DB[5][0] = 0.0;       DB[5][1] = D[5] * dz; DB[5][2] = D[5] * dy;
                for (int c = 0; c < nen; c++) {
                    int cd = c * dim;
                    double bx = GNx[0][c], by = GNx[1][c], bz = GNx[2][c];
                    for (int j = 0; j < 3; j++) {
                        Ke[(cd) * nd + rd + j] += bx * DB[0][j] + by * DB[3][j] + bz * DB[4][j];
                        Ke[(cd + 1) * nd + rd + j] += by * DB[1][j] + bx * DB[3][j] + bz * DB[5][j];
                        Ke[(cd + 2) * nd + rd + j] += bz * DB[2][j] + bx * DB[4][j] + by * DB[5][j];
                    }
                }
            }
        }
    }
}

/* _GradientStiffnessMatrixTerm_AssembleElement (GradientStiffnessMatrixTerm.c:160-250):
 * Ge[a*dim+i] += w * (-detJ) * GNx[i][a] * N_p  (N_p = 1 for P0, ConstantElementType.c:164-165). */
void orc_element_G(const ORC_Mesh *m, int e, int n, const double *xi, const double *w, double *Ge) {
    int dim = m->dim, nen = 1 << dim;
    for (int p = 0; p < n; p++) {
        double GNx[3][8], detJ;
        orc_shape_global_derivs(m, e, &xi[3 * p], &detJ, GNx);
        for (int a = 0; a < nen; a++)
            for (int i = 0; i < dim; i++) Ge[a * dim + i] += +w[p] * (-detJ) * (GNx[i][a] * 1.0);
    }
}

/* _BuoyancyForceTerm_AssembleElement (BuoyancyForceTerm.c:310-452) with alpha = beta = 0, no
 * gHat: f[a*dim + vert] += -1 * (detJ * w * rho * g) * N_a, vertical axis J. */
void orc_element_f(const ORC_Mesh *m, int e, int n, const double *xi, const double *w, const double *rho,
                   double gravity, double *fe) {
    int dim = m->dim, nen = 1 << dim;
    for (int p = 0; p < n; p++) {
        double GNx[3][8], detJ, N[8];
        orc_shape_global_derivs(m, e, &xi[3 * p], &detJ, GNx);
        orc_shape_functions(dim, &xi[3 * p], N);
        double factor = detJ * w[p] * 1.0 * (rho[p] * gravity);
        for (int a = 0; a < nen; a++) fe[a * dim + 1] += -1.0 * factor * N[a];
    }
}

/* HydrostaticTerm_Current_Heights and HydrostaticTerm_Density (HydrostaticTerm.c:267-330): the background
 * density at height y = coord[1] of the Gale-compatible hydrostatic term. */
static double orc_hydrostatic_density(const ORC_Buoyancy *b, double y) {
    double thick, boundary, height;
    if (b->hs_v != 0) {
        thick = (b->hs_height - b->hs_material_boundary) * b->hs_width / (b->hs_width + b->hs_v * b->hs_time);
        boundary = b->hs_material_boundary +
                   (b->hs_upper_density / b->hs_lower_density) * (b->hs_height - b->hs_material_boundary - thick);
    } else {
        thick = b->hs_height - b->hs_material_boundary;
        boundary = b->hs_material_boundary;
    }
    height = thick + boundary;
    double h = b->hs_height - y, T, density, alpha;
    if (h < 0) T = b->hs_T0;
    else T = b->hs_T0 + b->hs_linear_coefficient * h +
             b->hs_exponential_coefficient1 * (1 - exp(-b->hs_exponential_coefficient2 * h));
    if (y > height) density = alpha = 0;
    else if (y > boundary) {
        density = b->hs_upper_density;
        alpha = b->hs_upper_alpha;
    } else {
        density = b->hs_lower_density;
        alpha = b->hs_lower_alpha;
    }
    return density * (1 - alpha * T);
}

/* BuoyancyForceTerm.c:391-425: depth below the top of the box along VerticalAxis, temperature at xi,
 * rho = rho0 (1 - alpha (T - Tref)), times (1 + beta (depth - depthref)) when beta != 0, less the
 * hydrostatic background density. */
double orc_particle_density(const ORC_Mesh *m, const ORC_Material *mat, int e, const double *xi, const double *T,
                            const ORC_Buoyancy *b) {
    int dim = m->dim, nen = 1 << dim, nodes[8];
    double N[8], coord[3] = {0, 0, 0}, vx[3], temperature = 0.0;
    orc_element_nodes(m, e, nodes);
    orc_shape_functions(dim, xi, N);
    for (int a = 0; a < nen; a++) {
        orc_vertex(m, nodes[a], vx);
        for (int d = 0; d < dim; d++) coord[d] += N[a] * vx[d];
    }
    const int va = b ? b->vertical_axis : 1;
    const double depth = fabs(m->xmax[va] - coord[va]);
    if (T)
        for (int a = 0; a < nen; a++) temperature += N[a] * T[nodes[a]];
    double background = 0.0;
    if (b && b->hydrostatic) background = orc_hydrostatic_density(b, coord[1]);
    double density = mat->density * (1.0 - mat->alpha * (temperature - mat->ref_temperature));
    if (mat->beta != 0) density *= (1.0 + mat->beta * (depth - mat->ref_depth));
    return density - background;
}

void orc_eval_density(const ORC_Mesh *m, const ORC_Material *mats, int np, const int *cell_off, const double *xi,
                      const int *mat_of, const double *T, const ORC_Buoyancy *b, double *rho) {
    int nel = orc_nelements(m);
    (void)np;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nel; e++)
        for (int p = cell_off[e]; p < cell_off[e + 1]; p++)
            rho[p] = orc_particle_density(m, &mats[mat_of ? mat_of[p] : 0], e, &xi[3 * (size_t)p], T, b);
}

/* BuoyancyForceTerm.c:362-375 (adjust: cellLocalVolume 2^dim over the sum of the weights) and :435-448
 * (factor = detJ w adjust rho g; along gHat, or -1 along VerticalAxis). */
void orc_element_force(const ORC_Mesh *m, int e, int n, const double *xi, const double *w, const double *rho,
                       double gravity, const ORC_Buoyancy *b, double *fe) {
    int dim = m->dim, nen = 1 << dim;
    double adjust = 1.0;
    if (b && b->adjust) {
        double tw = 0.0;
        for (int p = 0; p < n; p++) tw += w[p];
        adjust = (double)(1 << dim) / tw;
    }
    const int va = b ? b->vertical_axis : 1;
    for (int p = 0; p < n; p++) {
        double GNx[3][8], detJ, N[8];
        orc_shape_global_derivs(m, e, &xi[3 * p], &detJ, GNx);
        orc_shape_functions(dim, &xi[3 * p], N);
        const double factor = detJ * w[p] * adjust * (rho[p] * gravity);
        for (int a = 0; a < nen; a++) {
            if (b && b->use_direction) {
                for (int d = 0; d < dim; d++) fe[a * dim + d] += b->direction[d] * factor * N[a];
            } else {
                fe[a * dim + va] += -1.0 * factor * N[a];
            }
        }
    }
}

/* _UzawaPreconditionerTerm_AssembleElement (UzawaPreconditionerTerm.c:195-212), P0: 1x1. */
double orc_element_uzawa(int dim, const double *Ke, const double *Ge) {
    int nd = (1 << dim) * dim;
    double s = 0.0;
    for (int v = 0; v < nd; v++) s += Ge[v] * (Ge[v] / Ke[v * nd + v]);
    return s;
}

void orc_assemble_elements(const ORC_Mesh *m, const int *cell_off, const double *xi, const double *w,
                           const double *eta, const double *rho, double gravity, double *Ke, double *Ge,
                           double *fe, double *approxS) {
    int dim = m->dim, nd = (1 << dim) * dim, nel = orc_nelements(m);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nel; e++) {
        int p0 = cell_off[e], n = cell_off[e + 1] - p0;
        double kloc[24 * 24], gloc[24];
        memset(kloc, 0, sizeof(kloc));
        memset(gloc, 0, sizeof(gloc));
        orc_element_K(m, e, n, &xi[3 * (size_t)p0], &w[p0], &eta[p0], kloc);
        orc_element_G(m, e, n, &xi[3 * (size_t)p0], &w[p0], gloc);
        if (Ke) memcpy(&Ke[(size_t)e * nd * nd], kloc, sizeof(double) * nd * nd);
        if (Ge) memcpy(&Ge[(size_t)e * nd], gloc, sizeof(double) * nd);
        if (approxS) approxS[e] = orc_element_uzawa(dim, kloc, gloc);
        if (fe) {
            memset(&fe[(size_t)e * nd], 0, sizeof(double) * nd);
            if (rho) orc_element_f(m, e, n, &xi[3 * (size_t)p0], &w[p0], &rho[p0], gravity, &fe[(size_t)e * nd]);
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * CSR helpers and global assembly (StiffnessMatrix.c:825-1017; eq numbers with BCs removed,
 * FeEquationNumber.c:186; BC lifting :900-951).
 * ---------------------------------------------------------------------------------------- */
void orc_csr_free(ORC_CSR *A) {
    free(A->rowptr);
    free(A->col);
    free(A->val);
    memset(A, 0, sizeof(*A));
}

void orc_csr_spmv(const ORC_CSR *A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A->n; i++) {
        double s = 0.0;
        for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++) s += A->val[k] * x[A->col[k]];
        y[i] = s;
    }
}

void orc_csr_spmv_t(const ORC_CSR *A, const double *x, double *y) {
    memset(y, 0, sizeof(double) * A->m);
    for (int i = 0; i < A->n; i++)
        for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++) y[A->col[k]] += A->val[k] * x[i];
}

int orc_eqnum(const ORC_Mesh *m, const unsigned char *bc_mask, int *eq) {
    int nn = orc_nnodes(m), dim = m->dim, c = 0;
    for (int n = 0; n < nn; n++)
        for (int d = 0; d < dim; d++) eq[n * dim + d] = (bc_mask && bc_mask[n * dim + d]) ? -1 : c++;
    return c;
}

static int csr_find(const ORC_CSR *A, int r, int c) {
    int lo = A->rowptr[r], hi = A->rowptr[r + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (A->col[mid] == c) return mid;
        if (A->col[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* Sparsity of the velocity-velocity block: nodes of the 3^dim neighbourhood. */
static void build_K_pattern(const ORC_Mesh *m, const int *eqmap, int ncols_full, ORC_CSR *K, int nrows) {
    int dim = m->dim, nn = orc_nnodes(m);
    int nx = m->nel[0] + 1, ny = m->nel[1] + 1, nz = dim == 3 ? m->nel[2] + 1 : 1;
    (void)ncols_full;
    K->n = K->m = nrows;
    K->rowptr = (int *)calloc((size_t)nrows + 1, sizeof(int));
    /* count */
    for (int n = 0; n < nn; n++) {
        int i = n % nx, j = (n / nx) % ny, k = n / (nx * ny), cnt = 0;
        for (int dk = -1; dk <= 1; dk++) for (int dj = -1; dj <= 1; dj++) for (int di = -1; di <= 1; di++) {
            if (dim == 2 && dk != 0) continue;
            int ii = i + di, jj = j + dj, kk = k + dk;
            if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) continue;
            int nb = ii + nx * (jj + ny * kk);
            for (int d = 0; d < dim; d++) if (eqmap[nb * dim + d] >= 0) cnt++;
        }
        for (int d = 0; d < dim; d++) if (eqmap[n * dim + d] >= 0) K->rowptr[eqmap[n * dim + d] + 1] = cnt;
    }
    for (int r = 0; r < nrows; r++) K->rowptr[r + 1] += K->rowptr[r];
    size_t nnz = (size_t)K->rowptr[nrows];
    K->col = (int *)malloc(nnz * sizeof(int));
    K->val = (double *)calloc(nnz, sizeof(double));
    for (int n = 0; n < nn; n++) {
        int i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
        int cols[81], cnt = 0;
        for (int dk = -1; dk <= 1; dk++) for (int dj = -1; dj <= 1; dj++) for (int di = -1; di <= 1; di++) {
            if (dim == 2 && dk != 0) continue;
            int ii = i + di, jj = j + dj, kk = k + dk;
            if (ii < 0 || jj < 0 || kk < 0 || ii >= nx || jj >= ny || kk >= nz) continue;
            int nb = ii + nx * (jj + ny * kk);
            for (int d = 0; d < dim; d++) if (eqmap[nb * dim + d] >= 0) cols[cnt++] = eqmap[nb * dim + d];
        }
        /* neighbourhood is visited in increasing node order, eq numbers increase with node: sorted */
        for (int d = 0; d < dim; d++) {
            int r = eqmap[n * dim + d];
            if (r < 0) continue;
            memcpy(&K->col[K->rowptr[r]], cols, sizeof(int) * cnt);
        }
    }
}

void orc_assemble_global(const ORC_Mesh *m, const int *eq, int neq, const unsigned char *bc_mask,
                         const double *bc_val, const double *Ke, const double *Ge, const double *fe,
                         ORC_CSR *K, ORC_CSR *G, double *f, double *h) {
    int dim = m->dim, nen = 1 << dim, nd = nen * dim, nel = orc_nelements(m);
    build_K_pattern(m, eq, neq, K, neq);
    /* G: row per free velocity dof, columns = adjacent elements (sorted) */
    G->n = neq;
    G->m = nel;
    G->rowptr = (int *)calloc((size_t)neq + 1, sizeof(int));
    for (int e = 0; e < nel; e++) {
        int nodes[8];
        orc_element_nodes(m, e, nodes);
        for (int a = 0; a < nen; a++)
            for (int i = 0; i < dim; i++) {
                int r = eq[nodes[a] * dim + i];
                if (r >= 0) G->rowptr[r + 1]++;
            }
    }
    for (int r = 0; r < neq; r++) G->rowptr[r + 1] += G->rowptr[r];
    G->col = (int *)malloc(sizeof(int) * (size_t)G->rowptr[neq]);
    G->val = (double *)calloc((size_t)G->rowptr[neq], sizeof(double));
    int *fill = (int *)calloc((size_t)neq, sizeof(int));
    memset(f, 0, sizeof(double) * neq);
    memset(h, 0, sizeof(double) * nel);
    for (int e = 0; e < nel; e++) { /* element order => G columns sorted */
        int nodes[8], ldof[24];
        orc_element_nodes(m, e, nodes);
        for (int a = 0; a < nen; a++)
            for (int i = 0; i < dim; i++) ldof[a * dim + i] = nodes[a] * dim + i;
        const double *ke = &Ke[(size_t)e * nd * nd];
        const double *ge = &Ge[(size_t)e * nd];
        for (int r = 0; r < nd; r++) {
            int R = eq[ldof[r]];
            if (R < 0) continue;
            double lift = 0.0;
            for (int c = 0; c < nd; c++) {
                int C = eq[ldof[c]];
                if (C >= 0) {
                    int pos = csr_find(K, R, C);
                    K->val[pos] += ke[r * nd + c];
                } else if (bc_mask && bc_mask[ldof[c]]) {
                    lift -= bc_val[ldof[c]] * ke[r * nd + c];
                }
            }
            f[R] += (fe ? fe[(size_t)e * nd + r] : 0.0) + lift;
            G->col[G->rowptr[R] + fill[R]] = e;
            G->val[G->rowptr[R] + fill[R]] = ge[r];
            fill[R]++;
        }
        for (int r = 0; r < nd; r++)
            if (eq[ldof[r]] < 0 && bc_mask && bc_mask[ldof[r]]) h[e] -= bc_val[ldof[r]] * ge[r];
    }
    free(fill);
}

void orc_assemble_K_full(const ORC_Mesh *m, const double *Ke, ORC_CSR *K) {
    int dim = m->dim, nn = orc_nnodes(m);
    int *eq = (int *)malloc(sizeof(int) * nn * dim);
    int neq = orc_eqnum(m, NULL, eq);
    int nel = orc_nelements(m), nen = 1 << dim, nd = nen * dim;
    build_K_pattern(m, eq, neq, K, neq);
    for (int e = 0; e < nel; e++) {
        int nodes[8];
        orc_element_nodes(m, e, nodes);
        for (int r = 0; r < nd; r++)
            for (int c = 0; c < nd; c++) {
                int R = nodes[r / dim] * dim + r % dim, C = nodes[c / dim] * dim + c % dim;
                K->val[csr_find(K, R, C)] += Ke[(size_t)e * nd * nd + r * nd + c];
            }
    }
    free(eq);
}

/* ------------------------------------------------------------------------------------------
 * Geometric multigrid (PETScMGSolver.c:459-716 with SROpGenerator.c:765-1083 operators):
 * P at the finest level maps all dofs of the next-coarser grid to the fine free dofs (BC rows
 * skipped, :1031); coarser P map all dofs to all dofs. Fine node at grid index i takes the
 * average of coarse nodes (i>>1)+{0,1} in each odd direction (:1007-1039).  Coarse operators
 * are Galerkin P^T A P (MatPtAP, :518-527).  Smoother: symmetric SOR (PETSc PCSOR default,
 * omega = 1), coarse solve: dense Cholesky (PETSc redundant LU role).
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    ORC_CSR A;     /* level operator */
    ORC_CSR P;     /* prolongation from the next coarser level (NULL on coarsest) */
    double *x, *b, *r;
    int n;
} MGLevel;

typedef struct {
    int nlev;
    MGLevel lev[16];
    double *Lc; /* dense Cholesky factor of coarsest */
    int nc;
    int smooth_its;
} MG;

static void build_P(const ORC_Mesh *fm, int dim, const int *feq, int nrows, ORC_CSR *P) {
    /* fine mesh fm (nodes), coarse mesh has nel/2 */
    int nxf = fm->nel[0] + 1, nyf = fm->nel[1] + 1, nzf = dim == 3 ? fm->nel[2] + 1 : 1;
    int nxc = fm->nel[0] / 2 + 1, nyc = fm->nel[1] / 2 + 1, nzc = dim == 3 ? fm->nel[2] / 2 + 1 : 1;
    int nnf = nxf * nyf * nzf;
    P->n = nrows;
    P->m = nxc * nyc * nzc * dim;
    P->rowptr = (int *)calloc((size_t)nrows + 1, sizeof(int));
    P->col = (int *)malloc(sizeof(int) * (size_t)nrows * 8);
    P->val = (double *)malloc(sizeof(double) * (size_t)nrows * 8);
    int nnz = 0;
    for (int n = 0; n < nnf; n++) {
        int idx[3] = {n % nxf, (n / nxf) % nyf, n / (nxf * nyf)};
        int noff[3] = {1, 1, 1};
        for (int d = 0; d < dim; d++) noff[d] = (idx[d] & 1) ? 2 : 1;
        int np = noff[0] * noff[1] * noff[2], cn[8], c = 0;
        for (int ok = 0; ok < noff[2]; ok++) for (int oj = 0; oj < noff[1]; oj++) for (int oi = 0; oi < noff[0]; oi++) {
            int ci = (idx[0] >> 1) + oi, cj = (idx[1] >> 1) + oj, ck = (idx[2] >> 1) + ok;
            cn[c++] = ci + nxc * (cj + nyc * ck);
        }
        for (int k = 0; k < dim; k++) {
            int row = feq ? feq[n * dim + k] : n * dim + k;
            if (row < 0) continue;
            /* rows are visited in increasing order */
            for (int j = 0; j < np; j++) {
                P->col[nnz] = cn[j] * dim + k;
                P->val[nnz] = 1.0 / (double)np;
                nnz++;
            }
            P->rowptr[row + 1] = nnz;
        }
    }
    for (int r = 0; r < nrows; r++)
        if (P->rowptr[r + 1] < P->rowptr[r]) P->rowptr[r + 1] = P->rowptr[r];
}

/* C = P^T A P (Gustavson with a dense accumulator) */
static void galerkin(const ORC_CSR *A, const ORC_CSR *P, ORC_CSR *C) {
    int nf = A->n, nc = P->m;
    /* transpose P */
    ORC_CSR Pt;
    Pt.n = nc; Pt.m = nf;
    Pt.rowptr = (int *)calloc((size_t)nc + 1, sizeof(int));
    for (int k = 0; k < P->rowptr[nf]; k++) Pt.rowptr[P->col[k] + 1]++;
    for (int i = 0; i < nc; i++) Pt.rowptr[i + 1] += Pt.rowptr[i];
    Pt.col = (int *)malloc(sizeof(int) * (size_t)Pt.rowptr[nc]);
    Pt.val = (double *)malloc(sizeof(double) * (size_t)Pt.rowptr[nc]);
    int *fill = (int *)calloc((size_t)nc, sizeof(int));
    for (int i = 0; i < nf; i++)
        for (int k = P->rowptr[i]; k < P->rowptr[i + 1]; k++) {
            int c = P->col[k], pos = Pt.rowptr[c] + fill[c]++;
            Pt.col[pos] = i;
            Pt.val[pos] = P->val[k];
        }
    free(fill);
    /* AP = A*P (nf x nc) */
    double *acc = (double *)calloc((size_t)nc, sizeof(double));
    int *mark = (int *)malloc(sizeof(int) * (size_t)nc), *list = (int *)malloc(sizeof(int) * (size_t)nc);
    for (int i = 0; i < nc; i++) mark[i] = -1;
    ORC_CSR AP;
    AP.n = nf; AP.m = nc;
    AP.rowptr = (int *)calloc((size_t)nf + 1, sizeof(int));
    size_t cap = (size_t)nf * 64 + 16, nnz = 0;
    AP.col = (int *)malloc(sizeof(int) * cap);
    AP.val = (double *)malloc(sizeof(double) * cap);
    for (int i = 0; i < nf; i++) {
        int nl = 0;
        for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++) {
            int j = A->col[k];
            for (int l = P->rowptr[j]; l < P->rowptr[j + 1]; l++) {
                int c = P->col[l];
                if (mark[c] != i) { mark[c] = i; list[nl++] = c; acc[c] = 0.0; }
                acc[c] += A->val[k] * P->val[l];
            }
        }
        if (nnz + nl > cap) {
            cap = 2 * cap + nl;
            AP.col = (int *)realloc(AP.col, sizeof(int) * cap);
            AP.val = (double *)realloc(AP.val, sizeof(double) * cap);
        }
        for (int t = 0; t < nl; t++) { AP.col[nnz] = list[t]; AP.val[nnz] = acc[list[t]]; nnz++; }
        AP.rowptr[i + 1] = (int)nnz;
    }
    /* C = Pt * AP (nc x nc), columns sorted per row */
    for (int i = 0; i < nc; i++) mark[i] = -1;
    C->n = C->m = nc;
    C->rowptr = (int *)calloc((size_t)nc + 1, sizeof(int));
    cap = (size_t)nc * 81 + 16; nnz = 0;
    C->col = (int *)malloc(sizeof(int) * cap);
    C->val = (double *)malloc(sizeof(double) * cap);
    for (int i = 0; i < nc; i++) {
        int nl = 0;
        for (int k = Pt.rowptr[i]; k < Pt.rowptr[i + 1]; k++) {
            int j = Pt.col[k];
            for (int l = AP.rowptr[j]; l < AP.rowptr[j + 1]; l++) {
                int c = AP.col[l];
                if (mark[c] != i) { mark[c] = i; list[nl++] = c; acc[c] = 0.0; }
                acc[c] += Pt.val[k] * AP.val[l];
            }
        }
        /* insertion sort of the (short) column list */
        for (int a = 1; a < nl; a++) {
            int v = list[a], b = a - 1;
            while (b >= 0 && list[b] > v) { list[b + 1] = list[b]; b--; }
            list[b + 1] = v;
        }
        if (nnz + nl > cap) {
            cap = 2 * cap + nl;
            C->col = (int *)realloc(C->col, sizeof(int) * cap);
            C->val = (double *)realloc(C->val, sizeof(double) * cap);
        }
        for (int t = 0; t < nl; t++) { C->col[nnz] = list[t]; C->val[nnz] = acc[list[t]]; nnz++; }
        C->rowptr[i + 1] = (int)nnz;
    }
    free(acc); free(mark); free(list);
    orc_csr_free(&AP);
    orc_csr_free(&Pt);
}

static double dot(int n, const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* symmetric Gauss-Seidel sweep on A x = b (PCSOR SOR_LOCAL_SYMMETRIC_SWEEP, omega = 1) */
static void ssor(const ORC_CSR *A, const double *b, double *x) {
    for (int pass = 0; pass < 2; pass++) {
        int beg = pass ? A->n - 1 : 0, end = pass ? -1 : A->n, st = pass ? -1 : 1;
        for (int i = beg; i != end; i += st) {
            double s = b[i], d = 0.0;
            for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; k++) {
                if (A->col[k] == i) d = A->val[k];
                else s -= A->val[k] * x[A->col[k]];
            }
            x[i] = s / d;
        }
    }
}

static void mg_vcycle(MG *mg, int l) {
    MGLevel *L = &mg->lev[l];
    if (l == mg->nlev - 1) { /* dense Cholesky solve */
        int n = mg->nc;
        double *y = L->x;
        memcpy(y, L->b, sizeof(double) * n);
        for (int i = 0; i < n; i++) {
            double s = y[i];
            for (int k = 0; k < i; k++) s -= mg->Lc[(size_t)i * n + k] * y[k];
            y[i] = s / mg->Lc[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < n; k++) s -= mg->Lc[(size_t)k * n + i] * y[k];
            y[i] = s / mg->Lc[(size_t)i * n + i];
        }
        return;
    }
    MGLevel *C = &mg->lev[l + 1];
    for (int s = 0; s < mg->smooth_its; s++) ssor(&L->A, L->b, L->x);
    orc_csr_spmv(&L->A, L->x, L->r);
    for (int i = 0; i < L->n; i++) L->r[i] = L->b[i] - L->r[i];
    orc_csr_spmv_t(&L->P, L->r, C->b);
    memset(C->x, 0, sizeof(double) * C->n);
    mg_vcycle(mg, l + 1);
    orc_csr_spmv(&L->P, C->x, L->r);
    for (int i = 0; i < L->n; i++) L->x[i] += L->r[i];
    for (int s = 0; s < mg->smooth_its; s++) ssor(&L->A, L->b, L->x);
}

static void mg_setup(MG *mg, const ORC_Mesh *m, const int *eq, const ORC_CSR *K, int levels, int smooth_its) {
    int dim = m->dim;
    memset(mg, 0, sizeof(*mg));
    mg->smooth_its = smooth_its;
    /* number of levels: halve while every axis stays even and the grid stays >= 2 elements */
    int maxlev = 1;
    {
        int n[3] = {m->nel[0], m->nel[1], dim == 3 ? m->nel[2] : 2};
        while (1) {
            int ok = 1;
            for (int d = 0; d < dim; d++) if (n[d] % 2 || n[d] / 2 < 2) ok = 0;
            int ndof = 1;
            for (int d = 0; d < dim; d++) ndof *= n[d] + 1;
            if (!ok || ndof * dim <= 600 || maxlev >= 15) break;
            for (int d = 0; d < dim; d++) n[d] /= 2;
            maxlev++;
        }
    }
    mg->nlev = (levels > 0 && levels < maxlev) ? levels : maxlev;
    ORC_Mesh lm = *m;
    mg->lev[0].A = *K; /* borrowed */
    mg->lev[0].n = K->n;
    for (int l = 0; l + 1 < mg->nlev; l++) {
        build_P(&lm, dim, l == 0 ? eq : NULL, mg->lev[l].n, &mg->lev[l].P);
        galerkin(&mg->lev[l].A, &mg->lev[l].P, &mg->lev[l + 1].A);
        mg->lev[l + 1].n = mg->lev[l + 1].A.n;
        for (int d = 0; d < dim; d++) lm.nel[d] /= 2;
    }
    for (int l = 0; l < mg->nlev; l++) {
        mg->lev[l].x = (double *)calloc(mg->lev[l].n, sizeof(double));
        mg->lev[l].b = (double *)calloc(mg->lev[l].n, sizeof(double));
        mg->lev[l].r = (double *)calloc(mg->lev[l].n, sizeof(double));
    }
    /* dense Cholesky of the coarsest operator */
    ORC_CSR *Ac = &mg->lev[mg->nlev - 1].A;
    int n = Ac->n;
    mg->nc = n;
    mg->Lc = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++)
        for (int k = Ac->rowptr[i]; k < Ac->rowptr[i + 1]; k++) mg->Lc[(size_t)i * n + Ac->col[k]] = Ac->val[k];
    for (int j = 0; j < n; j++) {
        double s = mg->Lc[(size_t)j * n + j];
        for (int k = 0; k < j; k++) s -= mg->Lc[(size_t)j * n + k] * mg->Lc[(size_t)j * n + k];
        double djj = sqrt(s);
        mg->Lc[(size_t)j * n + j] = djj;
        for (int i = j + 1; i < n; i++) {
            double t = mg->Lc[(size_t)i * n + j];
            for (int k = 0; k < j; k++) t -= mg->Lc[(size_t)i * n + k] * mg->Lc[(size_t)j * n + k];
            mg->Lc[(size_t)i * n + j] = t / djj;
        }
    }
}

static void mg_free(MG *mg) {
    for (int l = 0; l < mg->nlev; l++) {
        if (l > 0) orc_csr_free(&mg->lev[l].A);
        if (l + 1 < mg->nlev) orc_csr_free(&mg->lev[l].P);
        free(mg->lev[l].x); free(mg->lev[l].b); free(mg->lev[l].r);
    }
    free(mg->Lc);
}

/* PCG with one V-cycle as preconditioner; returns iterations.  x holds the initial guess. */
static int mg_pcg(MG *mg, const ORC_CSR *K, const double *b, double *x, double rtol, int maxit) {
    int n = K->n, it = 0;
    double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
    double *pp = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
    orc_csr_spmv(K, x, q);
    for (int i = 0; i < n; i++) r[i] = b[i] - q[i];
    double bn = sqrt(dot(n, b, b)), rn = sqrt(dot(n, r, r));
    if (bn == 0.0) bn = 1.0;
    double rz = 0.0;
    while (rn / bn > rtol && it < maxit) {
        memcpy(mg->lev[0].b, r, sizeof(double) * n);
        memset(mg->lev[0].x, 0, sizeof(double) * n);
        mg_vcycle(mg, 0);
        memcpy(z, mg->lev[0].x, sizeof(double) * n);
        double rz_new = dot(n, r, z);
        if (it == 0) memcpy(pp, z, sizeof(double) * n);
        else {
            double beta = rz_new / rz;
            for (int i = 0; i < n; i++) pp[i] = z[i] + beta * pp[i];
        }
        rz = rz_new;
        orc_csr_spmv(K, pp, q);
        double alpha = rz / dot(n, pp, q);
        for (int i = 0; i < n; i++) { x[i] += alpha * pp[i]; r[i] -= alpha * q[i]; }
        rn = sqrt(dot(n, r, r));
        it++;
    }
    free(r); free(z); free(pp); free(q);
    return it;
}

int orc_velocity_solve(const ORC_Mesh *m, const int *eq, int neq, const unsigned char *bc_mask,
                       const ORC_CSR *K, const double *b, double *u, const ORC_SolveOpts *opt, int *its,
                       double *t_setup, double *t_solve) {
    MG mg;
    (void)bc_mask; (void)neq;
    double t0 = now_s();
    mg_setup(&mg, m, eq, K, opt->mg_levels, opt->smooth_its > 0 ? opt->smooth_its : 1);
    double t1 = now_s();
    *its = mg_pcg(&mg, K, b, u, opt->inner_rtol, opt->max_inner > 0 ? opt->max_inner : 500);
    double t2 = now_s();
    if (t_setup) *t_setup = t1 - t0;
    if (t_solve) *t_solve = t2 - t1;
    mg_free(&mg);
    return 0;
}

double orc_stokes_residual(const ORC_CSR *K, const ORC_CSR *G, const double *f, const double *h, const double *u,
                           const double *p) {
    int n = K->n, ne = G->m;
    double *t = (double *)malloc(sizeof(double) * n), *s = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * ne);
    orc_csr_spmv(K, u, t);
    orc_csr_spmv(G, p, s);
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; i++) { double r = f[i] - t[i] - s[i]; num += r * r; den += f[i] * f[i]; }
    orc_csr_spmv_t(G, u, q);
    for (int e = 0; e < ne; e++) { double r = h[e] - q[e]; num += r * r; den += h[e] * h[e]; }
    free(t); free(s); free(q);
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

/* ------------------------------------------------------------------------------------------
 * BSSCR (auglag-driver-DGTGD.c:191-403, penalty 0): S = -G^T K^-1 G is solved as the positive
 * operator G^T K^-1 G with the sign of h_hat flipped (identical CG iterates).  h_hat = h - G^T K^-1 f
 * (:238-242); CG with the diagonal approxS as preconditioner (Q22 "uw": preconditioner.c:138-166,
 * a diagonal matrix under PETSc's default PC is Jacobi); constant-pressure null space projected out
 * (ksp_pressure_nullspace.c:121-176); back-solve u = K^-1 (f - G p) (:383-403).
 * ---------------------------------------------------------------------------------------- */
static void remove_const(int n, double *v) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += v[i];
    s /= n;
    for (int i = 0; i < n; i++) v[i] -= s;
}

int orc_bsscr_solve(const ORC_Mesh *m, const int *eq, int neq, const unsigned char *bc_mask, const ORC_CSR *K,
                    const ORC_CSR *G, const double *f, const double *h, const double *approxS,
                    const ORC_SolveOpts *opt, double *u, double *p, ORC_SolveStats *st) {
    int ne = G->m;
    MG mg;
    (void)bc_mask;
    double t0 = now_s();
    mg_setup(&mg, m, eq, K, opt->mg_levels, opt->smooth_its > 0 ? opt->smooth_its : 1);
    double t1 = now_s();
    int maxin = opt->max_inner > 0 ? opt->max_inner : 500, maxout = opt->max_outer > 0 ? opt->max_outer : 1000;
    memset(st, 0, sizeof(*st));
    double *tu = (double *)calloc(neq, sizeof(double)), *tv = (double *)malloc(sizeof(double) * neq);
    double *hh = (double *)malloc(sizeof(double) * ne), *r = (double *)malloc(sizeof(double) * ne);
    double *z = (double *)malloc(sizeof(double) * ne), *d = (double *)malloc(sizeof(double) * ne);
    double *q = (double *)malloc(sizeof(double) * ne);
    /* constant pressure in the null space of G?  ||G t||/||t|| < nstol ||D||_inf, nstol = 1e-7
     * (ksp_pressure_nullspace.c:139-158, BSSCR.c:391); -1 = that test, 0/1 = forced */
    int ns = opt->remove_nullspace;
    if (ns < 0) {
        double *one = (double *)malloc(sizeof(double) * ne), *g1 = (double *)malloc(sizeof(double) * neq);
        double *rs = (double *)calloc(ne, sizeof(double));
        for (int e = 0; e < ne; e++) one[e] = 1.0;
        orc_csr_spmv(G, one, g1);
        for (int i = 0; i < neq; i++)
            for (int k = G->rowptr[i]; k < G->rowptr[i + 1]; k++) rs[G->col[k]] += fabs(G->val[k]);
        double gnorm = 0.0;
        for (int e = 0; e < ne; e++) if (rs[e] > gnorm) gnorm = rs[e];
        double n1 = sqrt(dot(neq, g1, g1)) / sqrt((double)ne);
        ns = n1 < 1e-7 * gnorm;
        free(one); free(g1); free(rs);
    }
    /* h_hat (negated: solve G^T K^-1 G p = G^T K^-1 f - h) */
    st->inner_its_total += mg_pcg(&mg, K, f, tu, opt->inner_rtol, maxin);
    st->inner_solves++;
    orc_csr_spmv_t(G, tu, hh);
    for (int e = 0; e < ne; e++) hh[e] = hh[e] - h[e];
    if (ns) remove_const(ne, hh);
    memset(p, 0, sizeof(double) * ne);
    memcpy(r, hh, sizeof(double) * ne);
    double hn = sqrt(dot(ne, hh, hh)), rn = hn, rz = 0.0;
    if (hn == 0.0) hn = 1.0;
    int it = 0;
    while (rn / hn > opt->outer_rtol && it < maxout) {
        for (int e = 0; e < ne; e++) z[e] = r[e] / approxS[e];
        double rz_new = dot(ne, r, z);
        if (it == 0) memcpy(d, z, sizeof(double) * ne);
        else {
            double beta = rz_new / rz;
            for (int e = 0; e < ne; e++) d[e] = z[e] + beta * d[e];
        }
        rz = rz_new;
        /* q = G^T K^-1 G d */
        orc_csr_spmv(G, d, tv);
        memset(tu, 0, sizeof(double) * neq);
        st->inner_its_total += mg_pcg(&mg, K, tv, tu, opt->inner_rtol, maxin);
        st->inner_solves++;
        orc_csr_spmv_t(G, tu, q);
        double alpha = rz / dot(ne, d, q);
        for (int e = 0; e < ne; e++) { p[e] += alpha * d[e]; r[e] -= alpha * q[e]; }
        rn = sqrt(dot(ne, r, r));
        it++;
    }
    st->outer_its = it;
    if (ns) remove_const(ne, p);
    /* back solve u = K^-1 (f - G p) */
    orc_csr_spmv(G, p, tv);
    for (int i = 0; i < neq; i++) tv[i] = f[i] - tv[i];
    memset(u, 0, sizeof(double) * neq);
    st->inner_its_total += mg_pcg(&mg, K, tv, u, opt->inner_rtol, maxin);
    st->inner_solves++;
    double t2 = now_s();
    st->t_setup = t1 - t0;
    st->t_solve = t2 - t1;
    st->stokes_rel_residual = orc_stokes_residual(K, G, f, h, u, p);
    free(tu); free(tv); free(hh); free(r); free(z); free(d); free(q);
    mg_free(&mg);
    return 0;
}

/* ---- DVCWeights (PICellerator/Weights/src/DVCWeights.c) ----
 * The reference's own data structures: a boundary list per particle, one list of the cells just claimed, a
 * `done` flag per sub-cell for de-duplication.  Sub-grid and centres: _DVCWeights_ConstructGrid (:500-556,
 * centres accumulated from x0 - dx/2); seeds and rounds: _DVCWeights_CreateVoronoi (:668-713); claims:
 * _ClaimCells (:261-317) with _DistanceTest (:607-610); boundaries: _UpdateBchain (:395-446, neighbours in the
 * order S, N, E, W, U, D); centroids and weights: _GetCentroids (:203-231). */
static void dvc_axis_centres(int n, double *d, double *c) {
    const double dx = 2.0 / n;
    double x = -1.0 - dx / 2.0;
    for (int i = 0; i < n; i++) {
        x = x + dx;
        c[i] = x;
    }
    *d = dx;
}

int orc_dvc_cell(int dim, const int *res, int np, const double *pts, double *cen, double *w) {
    if (np < 1) return -1;
    const int nx = res[0], ny = res[1], nz = dim == 3 ? res[2] : 1, nxy = nx * ny, ncell = nxy * nz;
    double d[3] = {0.0, 0.0, 0.0}, ax[3][64];
    dvc_axis_centres(nx, &d[0], ax[0]);
    dvc_axis_centres(ny, &d[1], ax[1]);
    if (dim == 3) dvc_axis_centres(nz, &d[2], ax[2]);
    const int nn[3] = {nx, ny, nz};
    int *owner = (int *)malloc(sizeof(int) * ncell);
    char *done = (char *)calloc(ncell, 1);
    int *bound = (int *)malloc(sizeof(int) * (size_t)np * ncell); /* a boundary has at most ncell cells */
    int *nbound = (int *)calloc(np, sizeof(int));
    int *claimed = (int *)malloc(sizeof(int) * ncell);
    for (int c = 0; c < ncell; c++) owner[c] = -1;

#define DVC_CENTRE(c, a) ax[a][(a) == 0 ? (c) % nx : ((a) == 1 ? ((c) / nx) % ny : (c) / nxy)]
    /* the new boundary of p from the nclaimed cells in `claimed` */
#define DVC_UPDATE(p, nclaimed)                                                                           \
    do {                                                                                                  \
        int *b = bound + (size_t)(p) * ncell, cnt = 0;                                                    \
        for (int q = 0; q < (nclaimed); q++) {                                                            \
            const int c0 = claimed[q], i0 = c0 % nx, j0 = (c0 / nx) % ny, k0 = c0 / nxy;                  \
            const int nb[6] = {j0 > 0 ? c0 - nx : -2, j0 < ny - 1 ? c0 + nx : -2, i0 < nx - 1 ? c0 + 1 : -2, \
                               i0 > 0 ? c0 - 1 : -2, k0 < nz - 1 ? c0 + nxy : -2, k0 > 0 ? c0 - nxy : -2};  \
            for (int t = 0; t < 2 * dim; t++) {                                                           \
                const int c1 = nb[t];                                                                     \
                if (c1 != -2 && owner[c1] != (p) && !done[c1]) {                                          \
                    b[cnt++] = c1;                                                                        \
                    done[c1] = 1;                                                                         \
                }                                                                                         \
            }                                                                                             \
        }                                                                                                 \
        for (int q = 0; q < cnt; q++) done[b[q]] = 0;                                                     \
        nbound[p] = cnt;                                                                                  \
    } while (0)

    for (int p = 0; p < np; p++) {
        int idx[3] = {0, 0, 0};
        for (int a = 0; a < dim; a++) {
            const double t = (pts[p * dim + a] - (-1.0)) / d[a];
            idx[a] = t >= 0.0 ? (t < (double)nn[a] ? (int)t : nn[a] - 1) : 0;
        }
        const int seed = idx[0] + idx[1] * nx + idx[2] * nxy;
        owner[seed] = p;
        claimed[0] = seed;
        DVC_UPDATE(p, 1);
    }
    int total = 1;
    while (total != 0) {
        total = 0;
        for (int p = 0; p < np; p++) {
            const int *b = bound + (size_t)p * ncell;
            int nc = 0;
            for (int q = 0; q < nbound[p]; q++) {
                const int c = b[q], o = owner[c];
                int win = o == -1;
                if (!win && o != p) {
                    double s = 0.0;
                    for (int a = 0; a < dim; a++) {
                        const double x0 = DVC_CENTRE(c, a), x1 = pts[o * dim + a], x2 = pts[p * dim + a];
                        s = s + (x1 + x2 - x0 - x0) * (x1 - x2);
                    }
                    win = s > 0.0;
                }
                if (win) {
                    owner[c] = p;
                    claimed[nc++] = c;
                }
            }
            total += nc;
            DVC_UPDATE(p, nc);
        }
    }
#undef DVC_UPDATE
    const double vol = dim == 3 ? d[0] * d[1] * d[2] : d[0] * d[1];
    int *count = (int *)calloc(np, sizeof(int));
    for (int q = 0; q < np * dim; q++) cen[q] = 0.0;
    for (int c = 0; c < ncell; c++) {
        const int p = owner[c];
        for (int a = 0; a < dim; a++) cen[p * dim + a] += DVC_CENTRE(c, a);
        count[p]++;
    }
#undef DVC_CENTRE
    for (int p = 0; p < np; p++) {
        w[p] = count[p] * vol;
        for (int a = 0; a < dim; a++) cen[p * dim + a] = count[p] ? cen[p * dim + a] / count[p] : 0.0;
    }
    free(count);
    free(owner);
    free(done);
    free(bound);
    free(nbound);
    free(claimed);
    return 0;
}

int orc_dvc_swarm(int dim, const int *res, int nel, const int *cell_off, double *xi, double *w, int set_centroids) {
    for (int e = 0; e < nel; e++)
        if (cell_off[e + 1] <= cell_off[e]) return -(e + 1);
#pragma omp parallel for schedule(dynamic, 64)
    for (int e = 0; e < nel; e++) {
        const int a = cell_off[e], n = cell_off[e + 1] - a;
        double *pts = (double *)malloc(sizeof(double) * n * dim), *cen = (double *)malloc(sizeof(double) * n * dim);
        for (int p = 0; p < n; p++)
            for (int k = 0; k < dim; k++) pts[p * dim + k] = xi[(size_t)(a + p) * 3 + k];
        orc_dvc_cell(dim, res, n, pts, cen, w + a);
        if (set_centroids)
            for (int p = 0; p < n; p++)
                for (int k = 0; k < dim; k++) xi[(size_t)(a + p) * 3 + k] = cen[p * dim + k];
        free(pts);
        free(cen);
    }
    return 0;
}
