/*
 * ust_solver.c -- BSSCR Schur-complement solve on the GPU (C host driving the kernel ABI).
 *
 * Restates BSSCR_DRIVER_auglag (Solvers/KSPSolvers/src/BSSCR/auglag-driver-DGTGD.c:191-403) with
 * penalty 0: S = G^T K^-1 G (the reference's -D K^-1 G; CG iterates are identical), h_hat =
 * G^T K^-1 f - h, CG on S with the "uw" diagonal approxS (preconditioner.c:138-166), constant
 * pressure null space removed (ksp_pressure_nullspace.c:121-176), back-solve u = K^-1 (f - G p).
 * The A11 solves use CG preconditioned by a geometric multigrid V-cycle (PETScMGSolver.c:459-716
 * with the SROpGenerator transfers, SROpGenerator.c:1000-1041) whose Galerkin coarse operators
 * are formed exactly in the quadrature-weight representation; smoothing is Jacobi-Chebyshev
 * (a parallel smoother in place of the reference's sequential SOR) and the coarsest level is a
 * dense FP64 inverse.  Mixed precision runs the A11 solves in FP32 inside FP64 iterative
 * refinement; outer Krylov, residuals and the coarse solve stay FP64.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ust_internal.h"

#define S(ctx) ((ctx)->stream)

static double wall_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void dfree(void **p) {
    if (*p) ust_dev_free(*p);
    *p = NULL;
}
#define DFREE(x) dfree((void **)&(x))
#define DALLOC(ctx, ptr, bytes) CKD(ctx, ust_dev_alloc_zero(ctx, (void **)&(ptr), (bytes)))

static int grid_ndof(const ust_grid *g) {
    int n = 1;
    for (int d = 0; d < g->dim; d++) n *= g->nel[d] + 1;
    return n * g->dim;
}
static int grid_nel(const ust_grid *g) {
    int n = 1;
    for (int d = 0; d < g->dim; d++) n *= g->nel[d];
    return n;
}

void ust_graphs_reset(ust_context *ctx) {
    for (int k = 0; k < 3; k++) {
        if (ctx->pcg_exec[k]) ust_graph_destroy(ctx->pcg_exec[k]);
        ctx->pcg_exec[k] = NULL;
    }
}

void ust_levels_free(ust_context *ctx) {
    ust_graphs_reset(ctx);
    for (int l = 0; l < ctx->nlev; l++) {
        ust_level *L = &ctx->lev[l];
        if (l > 0) {
            DFREE(L->WK64);
            DFREE(L->mask);
            free(L->mask_h);
        }
        DFREE(L->WK32);
        DFREE(L->WK16);
        DFREE(L->invd64);
        DFREE(L->invd32);
        DFREE(L->x64); DFREE(L->b64); DFREE(L->ax64); DFREE(L->d64);
        DFREE(L->x32); DFREE(L->x32b); DFREE(L->b32); DFREE(L->ax32); DFREE(L->d32);
        memset(L, 0, sizeof(*L));
    }
    DFREE(ctx->Ainv);
    ctx->nlev = 0;
    ctx->levels_ready = 0;
    ctx->levels_prec32 = 0;
    ctx->levels_struct = 0;
}

/* ---------------- level operations (prec: 0 = FP64, 1 = FP32) ---------------- */
static int lev_apply(ust_context *ctx, int l, int f32, const void *x, void *y) {
    ust_level *L = &ctx->lev[l];
    if (f32 && L->WK16)
        CKD(ctx, ust_k_apply_K_f32_bf16w(&L->g, L->WK16, L->mask, (const float *)x, (float *)y, (float *)ctx->kwork,
                                         S(ctx)));
    else if (f32) CKD(ctx, ust_k_apply_K_f32(&L->g, L->WK32, L->mask, (const float *)x, (float *)y, (float *)ctx->kwork, S(ctx)));
    else CKD(ctx, ust_k_apply_K_f64(&L->g, L->WK64, L->mask, (const double *)x, (double *)y, (double *)ctx->kwork, S(ctx)));
    return UST_OK;
}

/* r = b - A x on level l (fused product and residual), r = 0 on the Dirichlet dofs */
static int lev_residual(ust_context *ctx, int l, int f32, const void *x, const void *b, void *r) {
    ust_level *L = &ctx->lev[l];
    if (f32 && L->WK16)
        CKD(ctx, ust_k_apply_K_residual_f32_bf16w(&L->g, L->WK16, L->mask, (const float *)x, (const float *)b,
                                                  (float *)r, (float *)ctx->kwork, S(ctx)));
    else if (f32)
        CKD(ctx, ust_k_apply_K_residual_f32(&L->g, L->WK32, L->mask, (const float *)x, (const float *)b, (float *)r,
                                            (float *)ctx->kwork, S(ctx)));
    else
        CKD(ctx, ust_k_apply_K_residual_f64(&L->g, L->WK64, L->mask, (const double *)x, (const double *)b,
                                            (double *)r, (double *)ctx->kwork, S(ctx)));
    return UST_OK;
}

/* Jacobi-Chebyshev smoothing of A x = b (PETScMGSolver.c:674-709 smoother role).  x_zero 1: x is not read
 * (taken as zero); the first step is then d = c2 D^-1 b, x = d, without a product.  x_zero 2: that first
 * step was already applied (by the restriction).  Every other step is one fused product + update pass.
 * *xp holds x; with alt (the level's second FP32 vector) the product passes run out of place, each writing
 * the other vector, and *xp / *alt are swapped after each (the FP32 tile-grid kernel needs it). */
static int lev_smooth(ust_context *ctx, int l, int f32, const void *b, void **xp, void **alt, int x_zero,
                      const void *xc) {
    ust_level *L = &ctx->lev[l];
    size_t n = (size_t)L->ndof;
    const int oop = f32 && alt && *alt && L->WK32 && !L->WK16;
    for (int k = x_zero == 2 ? 1 : 0; k < L->cheb_deg; k++) {
        void *x = *xp;
        if (k == 0 && xc) { /* prolongation of the coarse correction fused into the first step */
            /* a single post-smoothing step leaves its direction unwritten: no later step reads it */
            float *d32 = L->cheb_deg == 1 ? NULL : L->d32;
            if (oop)
                CKD(ctx, ust_k_apply_K_cheby_prolong_oop_f32(&L->g, L->WK32, L->mask, (const float *)x, (float *)*alt,
                                                             (const float *)b, L->invd32, d32, (float)L->c2[0],
                                                             (const float *)xc, (float *)ctx->kwork, S(ctx)));
            else if (f32 && L->WK16)
                CKD(ctx, ust_k_apply_K_cheby_prolong_f32_bf16w(&L->g, L->WK16, L->mask, (float *)x, (const float *)b,
                                                               L->invd32, L->d32, (float)L->c2[0], (const float *)xc,
                                                               (float *)ctx->kwork, S(ctx)));
            else if (f32)
                CKD(ctx, ust_k_apply_K_cheby_prolong_f32(&L->g, L->WK32, L->mask, (float *)x, (const float *)b,
                                                         L->invd32, d32, (float)L->c2[0], (const float *)xc,
                                                         (float *)ctx->kwork, S(ctx)));
            else
                CKD(ctx, ust_k_apply_K_cheby_prolong_f64(&L->g, L->WK64, L->mask, (double *)x, (const double *)b,
                                                         L->invd64, L->d64, L->c2[0], (const double *)xc,
                                                         (double *)ctx->kwork, S(ctx)));
        } else if (k == 0 && x_zero) {
            if (f32 && ctx->b64_src)
                CKD(ctx, ust_k_cheby_first_from_f64(ctx->b64_src, (float *)b, L->invd32, L->d32, (float *)x,
                                                    (float)L->c2[0], n, S(ctx)));
            else if (f32)
                CKD(ctx, ust_k_cheby_first_f32((const float *)b, L->invd32, L->d32, (float *)x, (float)L->c2[0], n,
                                               S(ctx)));
            else
                CKD(ctx, ust_k_cheby_first_f64((const double *)b, L->invd64, L->d64, (double *)x, L->c2[0], n,
                                               S(ctx)));
            continue; /* elementwise: in place whatever the level */
        } else if (oop) {
            CKD(ctx, ust_k_apply_K_cheby_oop_f32(&L->g, L->WK32, L->mask, (const float *)x, (float *)*alt,
                                                 (const float *)b, L->invd32, L->d32, (float)L->c1[k], (float)L->c2[k],
                                                 k == 0, (float *)ctx->kwork, S(ctx)));
        } else if (f32 && L->WK16) {
            CKD(ctx, ust_k_apply_K_cheby_f32_bf16w(&L->g, L->WK16, L->mask, (float *)x, (const float *)b, L->invd32,
                                                   L->d32, (float)L->c1[k], (float)L->c2[k], k == 0,
                                                   (float *)ctx->kwork, S(ctx)));
        } else if (f32) {
            CKD(ctx, ust_k_apply_K_cheby_f32(&L->g, L->WK32, L->mask, (float *)x, (const float *)b, L->invd32, L->d32,
                                             (float)L->c1[k], (float)L->c2[k], k == 0, (float *)ctx->kwork, S(ctx)));
        } else {
            CKD(ctx, ust_k_apply_K_cheby_f64(&L->g, L->WK64, L->mask, (double *)x, (const double *)b, L->invd64,
                                             L->d64, L->c1[k], L->c2[k], k == 0, (double *)ctx->kwork, S(ctx)));
        }
        if (oop) { /* the step wrote the other vector */
            *xp = *alt;
            *alt = x;
        }
    }
    return UST_OK;
}

/* one V-cycle: x = M^-1 b on level l (x is overwritten, not read).  first_done: the first smoothing step
 * from the zero guess was already applied to x (fused into the restriction that produced b).  *res: the
 * vector that holds the result -- x, or on an FP32 coarse level whose passes run out of place, possibly the
 * level's second vector x32b. */
static int vcycle_r(ust_context *ctx, int l, int f32, const void *b, void *x, int first_done, void **res) {
    ust_level *L = &ctx->lev[l];
    void *cur = x, *alt = (l > 0 && f32) ? (void *)L->x32b : NULL;
    int rc;
    *res = x;
    if (l == ctx->nlev - 1) {
        if (ctx->coarse_dense) {
            if (f32) CKD(ctx, ust_k_dense_matvec_f32(ctx->Ainv, ctx->nc, (const float *)b, (float *)x, S(ctx)));
            else CKD(ctx, ust_k_dense_matvec_f64(ctx->Ainv, ctx->nc, (const double *)b, (double *)x, S(ctx)));
            return UST_OK;
        }
        rc = lev_smooth(ctx, l, f32, b, &cur, &alt, first_done ? 2 : 1, NULL);
        *res = cur;
        return rc;
    }
    ust_level *C = &ctx->lev[l + 1];
    /* FP32, one pre-smoothing step from zero: the step rides on the residual pass */
    const int rfuse = f32 && !first_done && L->cheb_deg == 1 && L->WK32 && !L->WK16 && !ctx->b64_src;
    if (!rfuse && (rc = lev_smooth(ctx, l, f32, b, &cur, &alt, first_done ? 2 : 1, NULL))) return rc;
    ctx->b64_src = NULL;
    /* the coarse level starts with a smoothing step from zero unless it is the dense coarse solve */
    int fuse = C->cheb_deg > 0 && !(l + 1 == ctx->nlev - 1 && ctx->coarse_dense);
    /* 3D FP32: the prolongation of the correction rides on the first post-smoothing pass (FP64 keeps the
     * separate prolongation: the fused FP64 pass needs more registers than two waves per SIMD allow) */
    int pfuse = L->g.dim == 3 && L->cheb_deg > 0 && f32;
    void *xcr = NULL; /* the coarse level's result */
    if (f32) {
        /* the residual is zero on the Dirichlet dofs, so the restriction need not read the fine mask */
        if (rfuse)
            CKD(ctx, ust_k_apply_K_residual_first_f32(&L->g, L->WK32, L->mask, (const float *)b, L->invd32,
                                                      (float *)cur, L->ax32, (float)L->c2[0], (float *)ctx->kwork,
                                                      S(ctx)));
        else if ((rc = lev_residual(ctx, l, 1, cur, b, L->ax32)))
            return rc;
        if (fuse)
            CKD(ctx, ust_k_restrict_cheby_first_f32(&L->g, NULL, C->mask, L->ax32, C->b32, C->invd32, C->d32,
                                                    C->x32, (float)C->c2[0], S(ctx)));
        else
            CKD(ctx, ust_k_restrict_f32(&L->g, NULL, C->mask, L->ax32, C->b32, S(ctx)));
        if ((rc = vcycle_r(ctx, l + 1, 1, C->b32, C->x32, fuse, &xcr))) return rc;
        if (pfuse) {
            rc = lev_smooth(ctx, l, 1, b, &cur, &alt, 0, xcr);
            *res = cur;
            return rc;
        }
        CKD(ctx, ust_k_prolong_add_f32(&L->g, L->mask, (const float *)xcr, (float *)cur, S(ctx)));
    } else {
        if ((rc = lev_residual(ctx, l, 0, cur, b, L->ax64))) return rc;
        if (fuse)
            CKD(ctx, ust_k_restrict_cheby_first_f64(&L->g, NULL, C->mask, L->ax64, C->b64, C->invd64, C->d64,
                                                    C->x64, C->c2[0], S(ctx)));
        else
            CKD(ctx, ust_k_restrict_f64(&L->g, NULL, C->mask, L->ax64, C->b64, S(ctx)));
        if ((rc = vcycle_r(ctx, l + 1, 0, C->b64, C->x64, fuse, &xcr))) return rc;
        CKD(ctx, ust_k_prolong_add_f64(&L->g, L->mask, (const double *)xcr, (double *)cur, S(ctx)));
    }
    rc = lev_smooth(ctx, l, f32, b, &cur, &alt, 0, NULL);
    *res = cur;
    return rc;
}

/* the V-cycle on the finest level, whose passes run in place: the result is in x */
static int vcycle(ust_context *ctx, int l, int f32, const void *b, void *x, int first_done) {
    void *res = NULL;
    int rc = vcycle_r(ctx, l, f32, b, x, first_done, &res);
    if (rc == UST_OK && res != x) UST_FAIL(ctx, UST_EINVAL, "V-cycle result left in the level's second vector");
    return rc;
}

/* lambda_max of D^-1 A by power iteration (FP64) */
static int estimate_lmax(ust_context *ctx, int l, double *lmax) {
    ust_level *L = &ctx->lev[l];
    size_t n = (size_t)L->ndof;
    double *x = L->x64, *y = L->ax64, nrm, lam = 0.0;
    CKD(ctx, ust_k_fill_hash_f64(x, n, 12345 + l, S(ctx)));
    CKD(ctx, ust_k_dot_f64(x, x, n, ctx->partial, ctx->scal + SC_T0, S(ctx)));
    CKD(ctx, ust_read_scalar(ctx, SC_T0, &nrm));
    CKD(ctx, ust_k_axpby_f64(0.0, x, 1.0 / sqrt(nrm), x, n, S(ctx)));
    for (int it = 0; it < 15; it++) {
        int rc = lev_apply(ctx, l, 0, x, y);
        if (rc) return rc;
        CKD(ctx, ust_k_scale_mul_f64(L->invd64, y, y, n, S(ctx)));
        CKD(ctx, ust_k_dot_f64(y, y, n, ctx->partial, ctx->scal + SC_T0, S(ctx)));
        CKD(ctx, ust_read_scalar(ctx, SC_T0, &nrm));
        nrm = sqrt(nrm);
        if (!(nrm > 0.0)) break;
        lam = nrm;
        CKD(ctx, ust_k_axpby_f64(1.0 / nrm, y, 0.0, x, n, S(ctx)));
    }
    *lmax = lam;
    return UST_OK;
}

static int clamp_degree(int d, int dflt) {
    if (d <= 0) d = dflt;
    return d > UST_MAX_CHEBY ? UST_MAX_CHEBY : d;
}

static void cheby_coeffs(ust_level *L, double lmax, double ratio, int deg) {
    double hi = lmax, lo = hi / ratio;
    double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta, rho = 1.0 / sigma;
    L->lmax = lmax;
    L->cheb_deg = deg;
    L->c1[0] = 0.0;
    L->c2[0] = 1.0 / theta;
    for (int k = 1; k < deg; k++) {
        double rn = 1.0 / (2.0 * sigma - rho);
        L->c1[k] = rn * rho;
        L->c2[k] = 2.0 * rn / delta;
        rho = rn;
    }
}

/* Structure (allocations and BC masks of every level) is built once per mesh, BC set and precision;
 * the numeric part (Galerkin weights, FP32 copies, diagonals, coarse operator, Chebyshev bounds)
 * after every assembly. */
static int levels_structure(ust_context *ctx, const ust_solve_opts *o, int need32) {
    ust_levels_free(ctx);
    int dim = ctx->grid.dim, maxlev = o->mg_levels > 0 ? o->mg_levels : UST_MAX_LEVELS;
    if (maxlev > UST_MAX_LEVELS) maxlev = UST_MAX_LEVELS;
    ust_grid g = ctx->grid;
    int nlev = 1;
    /* the coarsest level takes the dense FP64 solve: coarsening stops at <= 1600 dofs, so a 128x128x64 grid
     * ends at 8x8x4 elements (1215 dofs, a 12 MB inverse whose product costs less than smoothing that level
     * and one more below it: 0.61 -> 0.55 s per solve at 128x128x64) */
    while (nlev < maxlev && grid_ndof(&g) > 1600) {
        int ok = 1;
        for (int d = 0; d < dim; d++)
            if (g.nel[d] % 2 || g.nel[d] < 2) ok = 0;
        if (!ok) break;
        for (int d = 0; d < dim; d++) {
            g.nel[d] /= 2;
            g.h[d] *= 2.0;
        }
        nlev++;
    }
    ctx->nlev = nlev;
    ctx->coarse_dense = grid_ndof(&g) <= 4096;
    g = ctx->grid;
    for (int l = 0; l < nlev; l++) {
        ust_level *L = &ctx->lev[l];
        L->g = g;
        L->ndof = grid_ndof(&g);
        L->nel = grid_nel(&g);
        L->nnodes = L->ndof / dim;
        if (l == 0) {
            L->WK64 = ctx->WK;
            L->mask = ctx->mask_d;
            L->mask_h = ctx->mask_h;
        } else {
            ust_level *F = &ctx->lev[l - 1];
            DALLOC(ctx, L->WK64, ust_k_weights_len(&L->g) * sizeof(double));
            /* a coarse node coincides with fine node (2I, 2J, 2K): it inherits that node's BC flags */
            L->mask_h = (unsigned char *)calloc((size_t)L->ndof, 1);
            if (!L->mask_h) UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
            int fnx = F->g.nel[0] + 1, fny = F->g.nel[1] + 1, cnx = g.nel[0] + 1, cny = g.nel[1] + 1;
            for (int n = 0; n < L->nnodes; n++) {
                int I = n % cnx, J = (n / cnx) % cny, K = n / (cnx * cny);
                int fn = 2 * I + fnx * (2 * J + fny * 2 * K);
                for (int c = 0; c < dim; c++) L->mask_h[n * dim + c] = F->mask_h[fn * dim + c];
            }
            DALLOC(ctx, L->mask, (size_t)L->ndof);
            CKD(ctx, ust_memcpy_h2d(L->mask, L->mask_h, (size_t)L->ndof, S(ctx)));
        }
        size_t n = (size_t)L->ndof;
        DALLOC(ctx, L->invd64, n * sizeof(double));
        DALLOC(ctx, L->x64, n * sizeof(double));
        DALLOC(ctx, L->b64, n * sizeof(double));
        DALLOC(ctx, L->ax64, n * sizeof(double));
        DALLOC(ctx, L->d64, n * sizeof(double));
        if (need32) {
            DALLOC(ctx, L->invd32, n * sizeof(float));
            if (o->vcycle_bf16) DALLOC(ctx, L->WK16, ust_k_weights_len(&L->g) * sizeof(unsigned short));
            else DALLOC(ctx, L->WK32, ust_k_weights_len(&L->g) * sizeof(float));
            DALLOC(ctx, L->x32, n * sizeof(float));
            /* coarse 3D levels smooth out of place with FP32 weights (the tile-grid kernel needs it) */
            if (l > 0 && dim == 3 && !o->vcycle_bf16) DALLOC(ctx, L->x32b, n * sizeof(float));
            DALLOC(ctx, L->b32, n * sizeof(float));
            DALLOC(ctx, L->ax32, n * sizeof(float));
            DALLOC(ctx, L->d32, n * sizeof(float));
        }
        if (l + 1 < nlev)
            for (int d = 0; d < dim; d++) {
                g.nel[d] /= 2;
                g.h[d] *= 2.0;
            }
    }
    /* the operator workspace serves every level: grow it if a coarse grid's tiling needs more */
    for (int l = 0; l < nlev; l++) {
        const size_t wl = ust_k_apply_work_len(&ctx->lev[l].g);
        if (wl > ctx->kwork_len) {
            DFREE(ctx->kwork);
            ctx->kwork_len = 0;
            DALLOC(ctx, ctx->kwork, wl * sizeof(double));
            ctx->kwork_len = wl;
        }
    }
    if (ctx->coarse_dense) {
        ctx->nc = ctx->lev[nlev - 1].ndof;
        /* the inverse plus the 4 nc pivot snapshots of its Gauss-Jordan construction */
        DALLOC(ctx, ctx->Ainv, ((size_t)ctx->nc * ctx->nc + 32 * (size_t)ctx->nc + 1024) * sizeof(double));
    }
    ctx->levels_prec32 = need32;
    ctx->levels_struct = 1;
    ctx->levels_deg = o->smoother_degree;
    ctx->levels_cdeg = o->coarse_degree;
    ctx->levels_ccdeg = o->coarsest_degree;
    ctx->levels_req = o->mg_levels;
    ctx->levels_w16 = need32 ? (o->vcycle_bf16 != 0) : -1;
    return UST_OK;
}

int ust_levels_setup(ust_context *ctx, const ust_solve_opts *o, int need32) {
    int rc;
    if (!ctx->levels_struct || (need32 && !ctx->levels_prec32) ||
        (need32 && ctx->levels_w16 != (o->vcycle_bf16 != 0)) || ctx->levels_deg != o->smoother_degree ||
        ctx->levels_cdeg != o->coarse_degree || ctx->levels_ccdeg != o->coarsest_degree ||
        ctx->levels_req != o->mg_levels) {
        if ((rc = levels_structure(ctx, o, need32 || ctx->levels_prec32))) return rc;
    }
    /* the FP32 velocity PCG applies K with FP32 weights on the finest level, beside the V-cycle's bfloat16 ones */
    if (need32 && o->inner_fp32 && !ctx->lev[0].WK32) {
        DALLOC(ctx, ctx->lev[0].WK32, ust_k_weights_len(&ctx->lev[0].g) * sizeof(float));
        ctx->levels_ready = 0;
    }
    if (ctx->levels_ready && ctx->levels_ratio == o->cheby_ratio) return UST_OK;
    ust_graphs_reset(ctx); /* captured smoother coefficients change with the operator */
    int nlev = ctx->nlev;
    for (int l = 0; l < nlev; l++) {
        ust_level *L = &ctx->lev[l];
        if (l > 0) CKD(ctx, ust_k_coarsen_weights(&ctx->lev[l - 1].g, ctx->lev[l - 1].WK64, L->WK64, S(ctx)));
        if (L->WK32) CKD(ctx, ust_k_copy_f64_to_f32(L->WK64, L->WK32, ust_k_weights_len(&L->g), S(ctx)));
        if (L->WK16) CKD(ctx, ust_k_copy_f64_to_bf16(L->WK64, L->WK16, ust_k_weights_len(&L->g), S(ctx)));
        CKD(ctx, ust_k_inv_diag(&L->g, L->WK64, L->mask, L->invd64, L->invd32, S(ctx)));
    }
    ust_level *Lc = &ctx->lev[nlev - 1];
    if (ctx->coarse_dense) { /* coarsest: dense FP64 inverse */
        CKD(ctx, ust_k_dense_assemble(&Lc->g, Lc->WK64, Lc->mask, ctx->Ainv, S(ctx)));
        CKD(ctx, ust_k_dense_spd_inverse(ctx->Ainv, ctx->nc, ctx->Ainv + (size_t)ctx->nc * ctx->nc, S(ctx)));
    }
    int deg = clamp_degree(o->smoother_degree, 2);
    int cdeg = clamp_degree(o->coarse_degree, deg), ccdeg = clamp_degree(o->coarsest_degree, UST_MAX_CHEBY);
    double ratio = o->cheby_ratio > 1.0 ? o->cheby_ratio : 12.0;
    for (int l = 0; l < nlev; l++) {
        if (l == nlev - 1 && ctx->coarse_dense) break;
        double lm;
        if ((rc = estimate_lmax(ctx, l, &lm))) return rc;
        if (l == nlev - 1)  /* grid too coarse-resistant for a dense solve: a long Chebyshev polynomial */
            cheby_coeffs(&ctx->lev[l], 1.1 * lm, 100.0, ccdeg);
        else
            cheby_coeffs(&ctx->lev[l], 1.1 * lm, ratio, l == 0 ? deg : cdeg);
    }
    ctx->levels_ready = 1;
    ctx->levels_ratio = o->cheby_ratio;
    CKD(ctx, ust_stream_sync(S(ctx)));
    return UST_OK;
}

static int ensure_work(ust_context *ctx, int need32) {
    size_t n = (size_t)ctx->ndof, ne = (size_t)ctx->nel;
    if (!ctx->vr) {
        DALLOC(ctx, ctx->vr, n * sizeof(double));
        DALLOC(ctx, ctx->vz, n * sizeof(double));
        DALLOC(ctx, ctx->vp, n * sizeof(double));
        DALLOC(ctx, ctx->vq, n * sizeof(double));
        DALLOC(ctx, ctx->vt, n * sizeof(double));
        DALLOC(ctx, ctx->vx, n * sizeof(double));
        DALLOC(ctx, ctx->va, n * sizeof(double));
        DALLOC(ctx, ctx->vb, n * sizeof(double));
        DALLOC(ctx, ctx->uprev, n * sizeof(double));
        DALLOC(ctx, ctx->pr, ne * sizeof(double));
        DALLOC(ctx, ctx->pz, ne * sizeof(double));
        DALLOC(ctx, ctx->pd, ne * sizeof(double));
        DALLOC(ctx, ctx->pq, ne * sizeof(double));
        DALLOC(ctx, ctx->ph, ne * sizeof(double));
        DALLOC(ctx, ctx->ppc, ne * sizeof(double));
        DALLOC(ctx, ctx->pprev, ne * sizeof(double));
    }
    if (need32 && !ctx->wr) {
        DALLOC(ctx, ctx->wr, n * sizeof(float));
        DALLOC(ctx, ctx->wz, n * sizeof(float));
        DALLOC(ctx, ctx->wp, n * sizeof(float));
        DALLOC(ctx, ctx->wq, n * sizeof(float));
        DALLOC(ctx, ctx->wx, n * sizeof(float));
    }
    return UST_OK;
}

static int dotv(ust_context *ctx, const double *a, const double *b, size_t n, int slot, double *out) {
    CKD(ctx, ust_k_dot_f64(a, b, n, ctx->partial, ctx->scal + slot, S(ctx)));
    if (out) CKD(ctx, ust_read_scalar(ctx, slot, out));
    return UST_OK;
}

/* One iteration of FP64 PCG on K x = b with one V-cycle as the preconditioner, on the context's work
 * vectors (x in vx).  pc32: the V-cycle runs in FP32 (FP32 weights, smoothers and transfers, FP64 coarse
 * solve) on the FP64 residual rounded to FP32; the Krylov vectors, K applies, dots and the stopping test
 * stay FP64, so each iteration is a refinement step that restores FP64 accuracy.  The scalars and the
 * stopping test live on the device (SC_CG block), so the sequence has no host round trip and is captured
 * once into a HIP graph. */
static int pcg_iteration(ust_context *ctx, int pc32) {
    size_t n = (size_t)ctx->ndof;
    double *r = ctx->vr, *z = ctx->vz, *p = ctx->vp, *q = ctx->vq, *sc = ctx->scal;
    int rc;
    if (pc32) {
        /* z stays FP32 (wz): the first smoothing step rounds r into b32 itself, r . z and p = z + beta p
         * read z as FP32 and widen it per element */
        ust_level *L0 = &ctx->lev[0];
        if (ctx->nlev > 1 && L0->cheb_deg > 0) {
            ctx->b64_src = r;
        } else {
            CKD(ctx, ust_k_copy_f64_to_f32(r, L0->b32, n, S(ctx)));
        }
        rc = vcycle(ctx, 0, 1, L0->b32, ctx->wz, 0);
        ctx->b64_src = NULL;
        if (rc) return rc;
        CKD(ctx, ust_k_dot_mixed_beta(r, ctx->wz, n, ctx->partial, sc + SC_RZ, sc + SC_RZOLD, sc + SC_CG, S(ctx)));
        CKD(ctx, ust_k_cg_p_guarded_zf(ctx->wz, p, sc + SC_CG, n, S(ctx)));
    } else {
        if ((rc = vcycle(ctx, 0, 0, r, z, 0))) return rc;
        CKD(ctx, ust_k_dot_beta_f64(r, z, n, ctx->partial, sc + SC_RZ, sc + SC_RZOLD, sc + SC_CG, S(ctx)));
        CKD(ctx, ust_k_cg_p_guarded_f64(z, p, sc + SC_CG, n, S(ctx)));
    }
    if ((rc = lev_apply(ctx, 0, 0, p, q))) return rc;
    CKD(ctx, ust_k_dot_f64(p, q, n, ctx->partial, sc + SC_PQ, S(ctx)));
    /* x, r update, r . r and the convergence check */
    CKD(ctx, ust_k_cg_xr_rr_guarded_f64(ctx->vx, r, p, q, sc + SC_RZ, sc + SC_PQ, sc + SC_CG, n, ctx->partial,
                                        sc + SC_RR, S(ctx)));
    return UST_OK;
}

static int pcg_graph(ust_context *ctx, int pc32) {
    if (ctx->pcg_exec[pc32]) return UST_OK;
    CKD(ctx, ust_graph_begin(S(ctx)));
    int rc = pcg_iteration(ctx, pc32);
    if (rc) {
        ust_graph_abort(S(ctx));
        return rc;
    }
    CKD(ctx, ust_graph_end(S(ctx), &ctx->pcg_exec[pc32]));
    return UST_OK;
}

/* PCG: x = K^-1 b to ||r|| <= rtol ||b|| from x = 0 (KSPSolve with KSPCG, StokesBlockKSPInterface.c:242).
 * The iterations run as graph launches; the host reads the device state after the number of iterations
 * the previous inner solve needed, then after every further one. */
static int pcg(ust_context *ctx, int pc32, const double *b, double *x, double rtol, int maxit, int *its) {
    size_t n = (size_t)ctx->ndof;
    double *sc = ctx->scal;
    int rc;
    CKD(ctx, ust_k_fill_f64(ctx->vx, 0.0, n, S(ctx)));
    CKD(ctx, ust_k_fill_f64(ctx->vp, 0.0, n, S(ctx)));
    CKD(ctx, ust_memcpy_d2d(ctx->vr, b, n * sizeof(double), S(ctx)));
    CKD(ctx, ust_k_dot_f64(b, b, n, ctx->partial, sc + SC_BB, S(ctx)));
    CKD(ctx, ust_k_cg_start(sc + SC_BB, rtol * rtol, sc + SC_CG, sc + SC_RZOLD, S(ctx)));
    if (ctx->use_graphs && (rc = pcg_graph(ctx, pc32))) return rc;
    int done = 0, first = ctx->pcg_hint > 2 ? ctx->pcg_hint - 1 : 1;
    double state = 0.0, count = 0.0;
    while (done < maxit) {
        int k = done == 0 ? first : 1;
        if (k > maxit - done) k = maxit - done;
        for (int j = 0; j < k; j++) {
            if (ctx->use_graphs) CKD(ctx, ust_graph_launch(ctx->pcg_exec[pc32], S(ctx)));
            else if ((rc = pcg_iteration(ctx, pc32))) return rc;
        }
        done += k;
        CKD(ctx, ust_memcpy_d2h(ctx->scal_h + SC_CG + 1, sc + SC_CG + 1, 2 * sizeof(double), S(ctx)));
        state = ctx->scal_h[SC_CG + 1];
        count = ctx->scal_h[SC_CG + 2];
        if (state != 0.0) break;
    }
    *its = (int)count;
    if (state == 2.0) UST_FAIL(ctx, UST_ENOTCONV, "A11 CG produced NaN (StokesBlockKSPInterface.c:32 check)");
    if (*its > 0) ctx->pcg_hint = *its;
    CKD(ctx, ust_memcpy_d2d(x, ctx->vx, n * sizeof(double), S(ctx)));
    return UST_OK;
}

/* One iteration of the FP32 velocity PCG (iterative refinement's inner solve): the V-cycle on the FP32 residual
 * wr, p = z + beta p, q = K p with the FP32 weights of the finest level, and the x (wx), r update; dots and r . r
 * accumulate in FP64 and the stopping test runs on the device, as in the FP64 PCG. */
static int pcg32_iteration(ust_context *ctx) {
    size_t n = (size_t)ctx->ndof;
    double *sc = ctx->scal;
    ust_level *L0 = &ctx->lev[0];
    int rc = vcycle(ctx, 0, 1, ctx->wr, ctx->wz, 0);
    if (rc) return rc;
    CKD(ctx, ust_k_dot_beta_f32(ctx->wr, ctx->wz, n, ctx->partial, sc + SC_RZ, sc + SC_RZOLD, sc + SC_CG, S(ctx)));
    CKD(ctx, ust_k_cg_p_guarded_f32(ctx->wz, ctx->wp, sc + SC_CG, n, S(ctx)));
    CKD(ctx, ust_k_apply_K_f32(&L0->g, L0->WK32, L0->mask, ctx->wp, ctx->wq, (float *)ctx->kwork, S(ctx)));
    CKD(ctx, ust_k_dot_f32(ctx->wp, ctx->wq, n, ctx->partial, sc + SC_PQ, S(ctx)));
    CKD(ctx, ust_k_cg_xr_rr_guarded_f32(ctx->wx, ctx->wr, ctx->wp, ctx->wq, sc + SC_RZ, sc + SC_PQ, sc + SC_CG, n,
                                        ctx->partial, sc + SC_RR, S(ctx)));
    return UST_OK;
}

/* FP32 PCG on K e = wr (wr at unit norm) to ||r|| <= rtol from e = 0; iterations as in pcg(), the launches
 * before the first check taken from the same refinement step of the previous solve (steps differ in length, so
 * one shared count would launch iterations past convergence) */
static int pcg32(ust_context *ctx, int step, double rtol, int maxit, int *its) {
    size_t n = (size_t)ctx->ndof;
    double *sc = ctx->scal;
    int rc;
    CKD(ctx, ust_k_fill_f32(ctx->wx, 0.0f, n, S(ctx)));
    CKD(ctx, ust_k_fill_f32(ctx->wp, 0.0f, n, S(ctx)));
    CKD(ctx, ust_k_dot_f32(ctx->wr, ctx->wr, n, ctx->partial, sc + SC_BB, S(ctx)));
    CKD(ctx, ust_k_cg_start(sc + SC_BB, rtol * rtol, sc + SC_CG, sc + SC_RZOLD, S(ctx)));
    if (ctx->use_graphs && !ctx->pcg_exec[2]) {
        CKD(ctx, ust_graph_begin(S(ctx)));
        if ((rc = pcg32_iteration(ctx))) {
            ust_graph_abort(S(ctx));
            return rc;
        }
        CKD(ctx, ust_graph_end(S(ctx), &ctx->pcg_exec[2]));
    }
    int *hint = &ctx->pcg32_hint[step < 16 ? step : 15];
    int done = 0, first = *hint > 2 ? *hint - 1 : 1;
    double state = 0.0, count = 0.0;
    while (done < maxit) {
        int k = done == 0 ? first : 1;
        if (k > maxit - done) k = maxit - done;
        for (int j = 0; j < k; j++) {
            if (ctx->use_graphs) CKD(ctx, ust_graph_launch(ctx->pcg_exec[2], S(ctx)));
            else if ((rc = pcg32_iteration(ctx))) return rc;
        }
        done += k;
        CKD(ctx, ust_memcpy_d2h(ctx->scal_h + SC_CG + 1, sc + SC_CG + 1, 2 * sizeof(double), S(ctx)));
        state = ctx->scal_h[SC_CG + 1];
        count = ctx->scal_h[SC_CG + 2];
        if (state != 0.0) break;
    }
    *its = (int)count;
    if (state == 2.0) UST_FAIL(ctx, UST_ENOTCONV, "A11 CG produced NaN (StokesBlockKSPInterface.c:32 check)");
    if (*its > 0) *hint = *its;
    return UST_OK;
}

/* FP32 correction solves to at most this relative reduction each (FP32 recurrences drift from the true residual
 * below it), and the FP64 residual is recomputed when the requested tolerance is tighter than IR_TRUST */
#ifndef IR_TAU_MIN
#define IR_TAU_MIN 1e-5
#endif
#ifndef IR_TRUST
#define IR_TRUST 1e-4
#endif
#define IR_MAX_STEPS 12

/* x = K^-1 b by iterative refinement: x accumulates in FP64 (in place, or in vt when x is b), r = b - K x in
 * FP64 (b itself before the first step, then vr); each step solves K e = r / ||r|| with the FP32 PCG and adds
 * ||r|| e */
static int pcg_refined(ust_context *ctx, const double *b, double *x, double rtol, int maxit, int *its,
                       int *steps) {
    size_t n = (size_t)ctx->ndof;
    double *sc = ctx->scal, bb, rr;
    double *acc = x == b ? ctx->vt : x;
    const double *r = b;
    int total = 0, k, rc;
    *steps = 0;
    CKD(ctx, ust_k_fill_f64(acc, 0.0, n, S(ctx)));
    CKD(ctx, ust_k_dot_f64(b, b, n, ctx->partial, sc + SC_T0, S(ctx)));
    CKD(ctx, ust_read_scalar(ctx, SC_T0, &bb));
    rr = bb;
    for (k = 0; k < IR_MAX_STEPS && total < maxit; k++) {
        if (!(rr > rtol * rtol * bb)) break;  /* converged (or b = 0) */
        const double rn = sqrt(rr), want = rtol * sqrt(bb) / rn;
        const double tau = want > IR_TAU_MIN ? want : IR_TAU_MIN;
        int it = 0;
        CKD(ctx, ust_k_copy_f64_to_f32_scaled(r, ctx->wr, 1.0 / rn, n, S(ctx)));
        if ((rc = pcg32(ctx, k, tau, maxit - total, &it))) return rc;
        total += it;
        CKD(ctx, ust_k_axpy_f32_to_f64(rn, ctx->wx, acc, n, S(ctx)));
        (*steps)++;
        if (want >= IR_TRUST) break;  /* a loose request: the FP32 recurrence is trusted */
        if ((rc = lev_apply(ctx, 0, 0, acc, ctx->vq))) return rc;
        CKD(ctx, ust_k_residual_f64(b, ctx->vq, ctx->vr, n, S(ctx)));
        r = ctx->vr;
        CKD(ctx, ust_k_dot_f64(ctx->vr, ctx->vr, n, ctx->partial, sc + SC_T0, S(ctx)));
        CKD(ctx, ust_read_scalar(ctx, SC_T0, &rr));
        if (!(rr == rr)) UST_FAIL(ctx, UST_ENOTCONV, "A11 refinement produced NaN");
    }
    *its = total;
    if (acc != x) CKD(ctx, ust_memcpy_d2d(x, acc, n * sizeof(double), S(ctx)));
    return UST_OK;
}

int ust_set_schur_scaling(ust_context *ctx, const float *log10_scale) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "set the mesh first");
    if (!log10_scale) UST_FAIL(ctx, UST_EINVAL, "no scaling given");
    if (!ctx->user_scale) DALLOC(ctx, ctx->user_scale, (size_t)ctx->nel * sizeof(float));
    CKD(ctx, ust_memcpy_h2d(ctx->user_scale, log10_scale, (size_t)ctx->nel * sizeof(float), S(ctx)));
    return UST_OK;
}

int ust_solver_prepare(ust_context *ctx, const ust_solve_opts *o) {
    int need32 = o->precision == UST_PREC_MIXED, rc;
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "assemble first");
    if ((rc = ust_levels_setup(ctx, o, need32))) return rc;
    return ensure_work(ctx, need32);
}

/* velocity solve K x = b (FP64 result) */
static int a11_solve(ust_context *ctx, const ust_solve_opts *o, const double *b, double *x, ust_solve_stats *st) {
    int its = 0, rc;
    st->inner_solves++;
    if (o->precision == UST_PREC_MIXED && o->inner_fp32) {
        int steps = 0;
        rc = pcg_refined(ctx, b, x, o->inner_rtol, o->max_inner, &its, &steps);
        st->inner_its += its;
        st->refinements += steps;
        return rc;
    }
    rc = pcg(ctx, o->precision == UST_PREC_MIXED, b, x, o->inner_rtol, o->max_inner, &its);
    st->inner_its += its;
    if (o->precision == UST_PREC_MIXED) st->refinements += its;
    return rc;
}

int ust_a11_solve_dev(ust_context *ctx, const ust_solve_opts *o, const double *b, double *x, ust_solve_stats *st) {
    return a11_solve(ctx, o, b, x, st);
}

/* q = G^T K^-1 G d  (the Schur complement apply of MatSchurComplement), inner solve to rtol */
static int schur_apply(ust_context *ctx, const ust_solve_opts *o, const double *d, double *q, double rtol,
                       ust_solve_stats *st) {
    CKD(ctx, ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, d, ctx->vb, S(ctx)));
    ust_solve_opts oi = *o;
    oi.inner_rtol = rtol;
    int rc = a11_solve(ctx, &oi, ctx->vb, ctx->uprev, st);
    if (rc) return rc;
    CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, ctx->uprev, q, 0, S(ctx)));
    return UST_OK;
}

static int remove_mean(ust_context *ctx, double *v) {
    CKD(ctx, ust_k_sum_f64(v, (size_t)ctx->nel, ctx->partial, ctx->scal + SC_SUM, S(ctx)));
    CKD(ctx, ust_k_remove_mean_f64(v, (size_t)ctx->nel, ctx->scal + SC_SUM, S(ctx)));
    return UST_OK;
}

/* preconditioned CG on the Schur complement from the current p.  *conv: the outer tolerance was met; *stalled:
 * the stall test (less than 10 % residual reduction over the last stall_window iterations) ended it. */
static int schur_cg(ust_context *ctx, const ust_solve_opts *o, int ns, const double *pc, int stall_window,
                    double hn, ust_solve_stats *st, int *stalled, int *conv) {
    size_t ne = (size_t)ctx->nel;
    double rn2;
    int rc;
    *stalled = 0;
    *conv = 0;
    /* residual norms of the last stall_window + 1 iterations (a ring: any iteration count keeps the test) */
    double *ring = NULL;
    if (stall_window > 0) {
        ring = (double *)malloc(sizeof(double) * (size_t)(stall_window + 1));
        if (!ring) UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    }
    rc = dotv(ctx, ctx->pr, ctx->pr, ne, SC_O_RR, &rn2);
    for (int it = 0; !rc && it < o->max_outer; it++) {
        double rn = sqrt(rn2);
        if (o->verbose) fprintf(stderr, "[ust] scr it %d |r|/|h| = %.3e\n", st->outer_its, rn / hn);
        if (rn <= o->outer_rtol * hn) {
            *conv = 1;
            break;
        }
        /* stalled: less than 10 % residual reduction over the last stall_window iterations (a slowly converging
         * learned run keeps going; a scaling that does not precondition is abandoned) */
        if (ring) {
            if (it >= stall_window && rn > 0.9 * ring[(it - stall_window) % (stall_window + 1)]) {
                *stalled = 1;
                break;
            }
            ring[it % (stall_window + 1)] = rn;
        }
        if ((rc = ust_k_div_f64(ctx->pr, pc, ctx->pz, ne, S(ctx)))) break;
        if ((rc = ust_k_dot_f64(ctx->pr, ctx->pz, ne, ctx->partial, ctx->scal + SC_O_RZ, S(ctx)))) break;
        if ((rc = ust_k_cg_update_p_f64(ctx->pz, ctx->pd, ctx->scal + SC_O_RZ, ctx->scal + SC_O_RZOLD, it == 0, ne,
                                        S(ctx))))
            break;
        if ((rc = ust_memcpy_d2d(ctx->scal + SC_O_RZOLD, ctx->scal + SC_O_RZ, sizeof(double), S(ctx)))) break;
        double rtol = o->inner_rtol;
        if (o->inner_relax && rn > 0.0) {
            rtol = o->inner_rtol * hn / rn;
            if (rtol > o->inner_relax_cap) rtol = o->inner_relax_cap;
            if (rtol < o->inner_rtol) rtol = o->inner_rtol;
        }
        if ((rc = schur_apply(ctx, o, ctx->pd, ctx->pq, rtol, st))) {
            free(ring);
            return rc;
        }
        if ((rc = ust_k_dot_f64(ctx->pd, ctx->pq, ne, ctx->partial, ctx->scal + SC_O_PQ, S(ctx)))) break;
        if ((rc = ust_k_cg_update_xr_f64(ctx->p, ctx->pr, ctx->pd, ctx->pq, ctx->scal + SC_O_RZOLD,
                                         ctx->scal + SC_O_PQ, ne, S(ctx))))
            break;
        if ((rc = dotv(ctx, ctx->pr, ctx->pr, ne, SC_O_RR, &rn2))) {
            free(ring);
            return rc;
        }
        st->outer_its++;
        if (!(rn2 == rn2)) {
            free(ring);
            UST_FAIL(ctx, UST_ENOTCONV, "Schur CG produced NaN");
        }
    }
    free(ring);
    if (rc) UST_FAIL(ctx, rc, "Schur CG step failed (%d): %s", rc, ust_dev_error_string());
    return UST_OK;
}

static int stokes_residual(ust_context *ctx, ust_solve_stats *st) {
    size_t n = (size_t)ctx->ndof, ne = (size_t)ctx->nel;
    double ff, hh, rm, rc2;
    int rc;
    /* momentum: f - K u - G p on free dofs (u here holds only the free part) */
    if ((rc = lev_apply(ctx, 0, 0, ctx->uprev, ctx->vt))) return rc;
    CKD(ctx, ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, ctx->p, ctx->vb, S(ctx)));
    CKD(ctx, ust_k_residual_f64(ctx->f, ctx->vt, ctx->va, n, S(ctx)));
    CKD(ctx, ust_k_axpy_f64(-1.0, ctx->vb, ctx->va, n, S(ctx)));
    CKD(ctx, dotv(ctx, ctx->va, ctx->va, n, SC_T0, &rm));
    CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, ctx->uprev, ctx->pq, 0, S(ctx)));
    CKD(ctx, ust_k_residual_f64(ctx->h, ctx->pq, ctx->pq, ne, S(ctx)));
    CKD(ctx, dotv(ctx, ctx->pq, ctx->pq, ne, SC_T0, &rc2));
    CKD(ctx, dotv(ctx, ctx->f, ctx->f, n, SC_T0, &ff));
    CKD(ctx, dotv(ctx, ctx->h, ctx->h, ne, SC_T0, &hh));
    double den = ff + hh > 0 ? ff + hh : 1.0;
    st->momentum_residual = sqrt(rm / den);
    st->continuity_residual = sqrt(rc2 / den);
    st->stokes_residual = sqrt((rm + rc2) / den);
    return UST_OK;
}

/* pressure null space, back-solve u = K^-1 (f - G p) and the FP64 Stokes residual.  With a Stokes-residual
 * bar the back-solve tolerance is also tied to it: the momentum residual it leaves is its relative residual
 * times ||f - G p||, which can exceed ||[f; h]|| by orders of magnitude when buoyancy and pressure gradient
 * nearly cancel (bs_scale: the fraction of the bar given to the back-solve). */
static int back_solve(ust_context *ctx, const ust_solve_opts *o, int ns, double bs_scale, ust_solve_stats *st) {
    size_t n = (size_t)ctx->ndof, ne = (size_t)ctx->nel;
    int rc;
    if (ns && (rc = remove_mean(ctx, ctx->p))) return rc;
    CKD(ctx, ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, ctx->p, ctx->vb, S(ctx)));
    CKD(ctx, ust_k_residual_f64(ctx->f, ctx->vb, ctx->vb, n, S(ctx)));
    ust_solve_opts ob = *o;
    if (o->stokes_rtol > 0.0) {
        double bb, ff, hh;
        CKD(ctx, dotv(ctx, ctx->vb, ctx->vb, n, SC_T1, &bb));
        CKD(ctx, dotv(ctx, ctx->f, ctx->f, n, SC_T1, &ff));
        CKD(ctx, dotv(ctx, ctx->h, ctx->h, ne, SC_T1, &hh));
        if (bb > 0.0) {
            double t = bs_scale * o->stokes_rtol * sqrt((ff + hh) / bb);
            if (t < ob.inner_rtol) ob.inner_rtol = t;
        }
    }
    if ((rc = a11_solve(ctx, &ob, ctx->vb, ctx->uprev, st))) return rc;
    return stokes_residual(ctx, st);
}

/* the diagonal Schur preconditioner of opts->schur_pc: approxS ("uw"), approxS scaled by the network or the
 * user's factors, or diag(G^T diag(K)^-1 G) ("gkgdiag", from the assembled diagonal of the finest level) */
static int schur_pc_vector(ust_context *ctx, const ust_solve_opts *o, const double **pc) {
    int rc;
    *pc = ctx->approxS;
    if (o->schur_pc == UST_SCHUR_UW) return UST_OK;
    if (o->schur_pc == UST_SCHUR_LEARNED) {
        if ((rc = ust_mlp_schur_pc(ctx, ctx->ppc))) return rc;
    } else if (o->schur_pc == UST_SCHUR_SCALED) {
        if (!ctx->user_scale) UST_FAIL(ctx, UST_ESTATE, "UST_SCHUR_SCALED needs ust_set_schur_scaling");
        CKD(ctx, ust_k_apply_scaling(ctx->approxS, ctx->user_scale, ctx->ppc, ctx->nel, S(ctx)));
    } else if (o->schur_pc == UST_SCHUR_GKGDIAG) {
        CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, ctx->lev[0].invd64, ctx->ppc, 3, S(ctx)));
    } else {
        UST_FAIL(ctx, UST_EINVAL, "unknown Schur preconditioner %d", o->schur_pc);
    }
    *pc = ctx->ppc;
    return UST_OK;
}

int ust_get_schur_preconditioner(ust_context *ctx, const ust_solve_opts *opts, double *pc_host) {
    ust_solve_opts o;
    if (opts) o = *opts;
    else ust_solve_opts_default(&o);
    int rc;
    if ((rc = ust_solver_prepare(ctx, &o))) return rc;
    const double *pc;
    if ((rc = schur_pc_vector(ctx, &o, &pc))) return rc;
    CKD(ctx, ust_memcpy_d2h(pc_host, pc, (size_t)ctx->nel * sizeof(double), S(ctx)));
    return UST_OK;
}

int ust_solve(ust_context *ctx, const ust_solve_opts *opts, ust_solve_stats *st) {
    ust_solve_opts o;
    if (opts) o = *opts;
    else ust_solve_opts_default(&o);
    memset(st, 0, sizeof(*st));
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "ust_assemble must run before ust_solve");
    int need32 = o.precision == UST_PREC_MIXED, rc;
    size_t n = (size_t)ctx->ndof, ne = (size_t)ctx->nel;
    double t0 = wall_ms();
    if ((rc = ust_levels_setup(ctx, &o, need32))) return rc;
    if ((rc = ensure_work(ctx, need32))) return rc;
    CKD(ctx, ust_stream_sync(S(ctx)));
    double t1 = wall_ms();
    st->mg_levels = ctx->nlev;
    int ns = o.remove_nullspace < 0 ? ctx->nullspace : o.remove_nullspace;
    st->nullspace_removed = ns;
    st->nullspace_ratio = ctx->ns_ratio;
    /* h_hat = G^T K^-1 f - h */
    if ((rc = a11_solve(ctx, &o, ctx->f, ctx->uprev, st))) return rc;
    CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, ctx->uprev, ctx->ph, 0, S(ctx)));
    CKD(ctx, ust_k_axpy_f64(-1.0, ctx->h, ctx->ph, ne, S(ctx)));
    if (ns && (rc = remove_mean(ctx, ctx->ph))) return rc;
    double hh;
    CKD(ctx, dotv(ctx, ctx->ph, ctx->ph, ne, SC_T0, &hh));
    double hn = sqrt(hh) > 0 ? sqrt(hh) : 1.0;
    /* h_hat tiny against ||G||_inf: the Schur solve runs on h_hat / hnorm and the pressure is scaled back
     * (auglag-driver-DGTGD.c:333-340, 356-358) */
    double hscale = 1.0;
    {
        double gnorm;
        CKD(ctx, ust_k_abs_G_rowsum(&ctx->grid, ctx->WG, ctx->mask_d, ctx->tmpv, S(ctx)));
        CKD(ctx, ust_k_absmax_f64(ctx->tmpv, n, ctx->partial, ctx->scal + SC_T0, S(ctx)));
        CKD(ctx, ust_read_scalar(ctx, SC_T0, &gnorm));
        const double hnorm = gnorm > 0.0 ? sqrt(hh) / gnorm : 0.0;
        if (hnorm < 1e-6 && hnorm > 1e-20) {
            hscale = hnorm;
            CKD(ctx, ust_k_axpby_f64(0.0, ctx->ph, 1.0 / hscale, ctx->ph, ne, S(ctx)));
            hn /= hscale;
            st->hhat_rescaled = 1;
        }
    }
    CKD(ctx, ust_k_fill_f64(ctx->p, 0.0, ne, S(ctx)));
    CKD(ctx, ust_memcpy_d2d(ctx->pr, ctx->ph, ne * sizeof(double), S(ctx)));
    int stalled = 0, conv = 0;
    const double *pc = ctx->approxS;
    const double *pc_last = ctx->approxS; /* the preconditioner the Schur CG finished with */
    if (o.schur_pc == UST_SCHUR_GKGDIAG) {
        if ((rc = schur_pc_vector(ctx, &o, &pc))) return rc;
        pc_last = pc;
        if ((rc = schur_cg(ctx, &o, ns, pc, 0, hn, st, &stalled, &conv))) return rc;
    } else if (o.schur_pc == UST_SCHUR_LEARNED || o.schur_pc == UST_SCHUR_SCALED) {
        if ((rc = schur_pc_vector(ctx, &o, &pc))) return rc;
        pc_last = pc;
        if ((rc = schur_cg(ctx, &o, ns, pc, o.stall_window, hn, st, &stalled, &conv))) return rc;
        if (stalled) {
            pc_last = ctx->approxS;
            /* plain BSSCR takes over from the current iterate: r = h_hat - S p */
            st->fallback_used = 1;
            if ((rc = schur_apply(ctx, &o, ctx->p, ctx->pq, o.inner_rtol, st))) return rc;
            CKD(ctx, ust_k_sub_f64(ctx->ph, ctx->pq, ctx->pr, ne, S(ctx)));
            if ((rc = schur_cg(ctx, &o, ns, ctx->approxS, 0, hn, st, &stalled, &conv))) return rc;
        }
    } else {
        if ((rc = schur_cg(ctx, &o, ns, pc, 0, hn, st, &stalled, &conv))) return rc;
    }
    if (hscale != 1.0) {
        CKD(ctx, ust_k_axpby_f64(0.0, ctx->ph, hscale, ctx->ph, ne, S(ctx)));
        CKD(ctx, ust_k_axpby_f64(0.0, ctx->p, hscale, ctx->p, ne, S(ctx)));
        hn *= hscale;
    }
    double bs_scale = 0.3;
    if ((rc = back_solve(ctx, &o, ns, bs_scale, st))) return rc;
    for (int pass = 1; o.stokes_rtol > 0.0 && st->stokes_residual > o.stokes_rtol && pass <= 3; pass++) {
        st->restarts++;
        if (st->continuity_residual > 0.5 * o.stokes_rtol) {
            /* inexact inner solves let the recursively updated Schur residual drift from the true one:
             * restart from the current pressure on r = h_hat - S p with a tighter outer tolerance */
            o.outer_rtol *= 0.1;
            if ((rc = schur_apply(ctx, &o, ctx->p, ctx->pq, o.inner_rtol, st))) return rc;
            CKD(ctx, ust_k_sub_f64(ctx->ph, ctx->pq, ctx->pr, ne, S(ctx)));
            if ((rc = schur_cg(ctx, &o, ns, pc_last, 0, hn, st, &stalled, &conv))) return rc;
        }
        bs_scale *= 0.1;
        if ((rc = back_solve(ctx, &o, ns, bs_scale, st))) return rc;
    }
    /* converged: the Stokes-residual bar when one is set, else the outer tolerance of the last Schur CG (max_outer
     * reached without it, as KSP's iteration limit, returns UST_OK with converged = 0) */
    st->converged = o.stokes_rtol > 0.0 ? st->stokes_residual <= o.stokes_rtol : conv;
    double t2 = wall_ms();
    /* nodal solution with BC values (rheology input of the next assembly) */
    CKD(ctx, ust_memcpy_d2d(ctx->u, ctx->uprev, n * sizeof(double), S(ctx)));
    CKD(ctx, ust_k_axpy_f64(1.0, ctx->bcval_d, ctx->u, n, S(ctx)));
    CKD(ctx, ust_stream_sync(S(ctx)));
    st->t_setup_ms = t1 - t0;
    st->t_solve_ms = t2 - t1;
    return UST_OK;
}

void ust_uzawa_opts_default(ust_uzawa_opts *o) {
    /* Stokes_SLE_UzawaSolver.c:258-262 */
    o->tolerance = 1e-5;
    o->max_iterations = 1000;
    o->min_iterations = 1;
    o->use_absolute_tolerance = 0;
    o->preconditioner = 1;
}

/* Stokes_SLE_UzawaSolver.c:407-983 (no compressibility matrix, D = G^T): steps 1-5 of the reference. */
int ust_solve_uzawa(ust_context *ctx, const ust_uzawa_opts *uzawa, const ust_solve_opts *vel, ust_solve_stats *st) {
    ust_uzawa_opts uo;
    ust_solve_opts o;
    if (uzawa) uo = *uzawa;
    else ust_uzawa_opts_default(&uo);
    if (vel) o = *vel;
    else ust_solve_opts_default(&o);
    memset(st, 0, sizeof(*st));
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "ust_assemble must run before ust_solve_uzawa");
    if (!(uo.tolerance > 0.0) || uo.max_iterations < 1) UST_FAIL(ctx, UST_EINVAL, "bad Uzawa tolerance or iterations");
    int need32 = o.precision == UST_PREC_MIXED, rc;
    size_t n = (size_t)ctx->ndof, ne = (size_t)ctx->nel;
    double t0 = wall_ms();
    if ((rc = ust_levels_setup(ctx, &o, need32))) return rc;
    if ((rc = ensure_work(ctx, need32))) return rc;
    double t1 = wall_ms();
    st->mg_levels = ctx->nlev;
    const double scale = 1.0 / sqrt((double)ne); /* qReciprocalGlobalProblemScale */
    double ff;
    CKD(ctx, dotv(ctx, ctx->f, ctx->f, n, SC_T0, &ff));
    if (sqrt(ff) / sqrt((double)n) <= 1e-99) /* Stokes_SLE_UzawaSolver.c:493-511 */
        UST_FAIL(ctx, UST_EINVAL, "the momentum force vector is zero");
    double *u = ctx->uprev, *q = ctx->p, *r = ctx->pr, *z = ctx->pz, *sdir = ctx->pd, *qt = ctx->pq;
    double *vstar = ctx->va, *ft = ctx->vb;
    /* step 1: the scale of the transformed right-hand side, G^T K^-1 f - h (velocity solve at tolerance) */
    o.inner_rtol = uo.tolerance;
    if ((rc = a11_solve(ctx, &o, ctx->f, vstar, st))) return rc;
    CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, vstar, qt, 0, S(ctx)));
    CKD(ctx, ust_k_axpy_f64(-1.0, ctx->h, qt, ne, S(ctx)));
    double qq;
    CKD(ctx, dotv(ctx, qt, qt, ne, SC_T0, &qq));
    const double rhs_scale = sqrt(qq) * scale;
    if (!isfinite(rhs_scale)) UST_FAIL(ctx, UST_ENOTCONV, "Uzawa right-hand side scale is not finite");
    /* step 2: all further velocity solves at 0.1 tolerance */
    o.inner_rtol = 0.1 * uo.tolerance;
    /* step 3: u0 = K^-1 (f - G q0), r = G^T u0 - h */
    CKD(ctx, ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, q, ft, S(ctx)));
    CKD(ctx, ust_k_residual_f64(ctx->f, ft, ft, n, S(ctx)));
    if ((rc = a11_solve(ctx, &o, ft, u, st))) return rc;
    CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, u, r, 0, S(ctx)));
    CKD(ctx, ust_k_axpy_f64(-1.0, ctx->h, r, ne, S(ctx)));
    /* step 4: preconditioned CG on the pressure */
    int it = 0;
    double zr_prev = 1.0, abs_res = 0.0, rel_res = 0.0;
    do {
        if (uo.preconditioner) CKD(ctx, ust_k_div_f64(r, ctx->approxS, z, ne, S(ctx)));
        else CKD(ctx, ust_memcpy_d2d(z, r, ne * sizeof(double), S(ctx)));
        double zr, sks;
        CKD(ctx, dotv(ctx, z, r, ne, SC_T0, &zr));
        if (it == 0) CKD(ctx, ust_memcpy_d2d(sdir, z, ne * sizeof(double), S(ctx)));
        else CKD(ctx, ust_k_axpby_f64(1.0, z, zr / zr_prev, sdir, ne, S(ctx)));
        /* K v* = G s */
        CKD(ctx, ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, sdir, ft, S(ctx)));
        if ((rc = a11_solve(ctx, &o, ft, vstar, st))) return rc;
        CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, vstar, qt, 0, S(ctx)));
        CKD(ctx, dotv(ctx, sdir, qt, ne, SC_T0, &sks));
        const double alpha = zr / sks;
        if (!isfinite(zr) || !isfinite(sks) || !isfinite(alpha))
            UST_FAIL(ctx, UST_ENOTCONV, "Uzawa step has an illegal value (Stokes_SLE_UzawaSolver.c:829-833)");
        CKD(ctx, ust_k_axpy_f64(alpha, sdir, q, ne, S(ctx)));
        CKD(ctx, ust_k_axpy_f64(-alpha, vstar, u, n, S(ctx)));
        CKD(ctx, ust_k_axpy_f64(-alpha, qt, r, ne, S(ctx)));
        zr_prev = zr;
        double rr;
        CKD(ctx, dotv(ctx, r, r, ne, SC_T0, &rr));
        abs_res = sqrt(rr) * scale;
        rel_res = abs_res / rhs_scale;
        if (o.verbose) fprintf(stderr, "[ust] uzawa %d abs %.6e rel %.6e\n", it, abs_res, rel_res);
        if (!isfinite(abs_res)) UST_FAIL(ctx, UST_ENOTCONV, "Uzawa residual is not finite");
        if (it >= uo.max_iterations) {
            st->outer_its = it;
            UST_FAIL(ctx, UST_ENOTCONV, "Uzawa reached maxIterations without converging");
        }
        it++;
    } while ((uo.use_absolute_tolerance ? abs_res : rel_res) > uo.tolerance || it < uo.min_iterations);
    st->outer_its = it;
    st->converged = 1;
    /* step 5: residuals, then the nodal velocity with its BC values */
    if ((rc = stokes_residual(ctx, st))) return rc;
    double t2 = wall_ms();
    CKD(ctx, ust_memcpy_d2d(ctx->u, u, n * sizeof(double), S(ctx)));
    CKD(ctx, ust_k_axpy_f64(1.0, ctx->bcval_d, ctx->u, n, S(ctx)));
    CKD(ctx, ust_stream_sync(S(ctx)));
    st->t_setup_ms = t1 - t0;
    st->t_solve_ms = t2 - t1;
    return UST_OK;
}

int ust_velocity_solve(ust_context *ctx, const ust_solve_opts *opts, const double *b, double *u, int *its) {
    ust_solve_opts o;
    if (opts) o = *opts;
    else ust_solve_opts_default(&o);
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "ust_assemble must run first");
    int need32 = o.precision == UST_PREC_MIXED, rc;
    size_t n = (size_t)ctx->ndof;
    if ((rc = ust_levels_setup(ctx, &o, need32))) return rc;
    if ((rc = ensure_work(ctx, need32))) return rc;
    ust_solve_stats st;
    memset(&st, 0, sizeof(st));
    CKD(ctx, ust_memcpy_h2d(ctx->vb, b, n * sizeof(double), S(ctx)));
    /* the right-hand side lives on the free dofs */
    CKD(ctx, ust_k_mask_zero_f64(ctx->vb, ctx->mask_d, n, S(ctx)));
    if ((rc = a11_solve(ctx, &o, ctx->vb, ctx->uprev, &st))) return rc;
    CKD(ctx, ust_memcpy_d2h(u, ctx->uprev, n * sizeof(double), S(ctx)));
    if (its) *its = st.inner_its;
    return UST_OK;
}

int ust_picard(ust_context *ctx, double gravity, double tol, int max_its, const ust_solve_opts *opts, int *its_done,
               double *final_change, ust_solve_stats *last) {
    size_t n = (size_t)ctx->ndof;
    int rc;
    double change = 1.0;
    int it;
    if (!ctx->upic) DALLOC(ctx, ctx->upic, n * sizeof(double));
    for (it = 0; it < max_its; it++) {
        if (it > 0) CKD(ctx, ust_memcpy_d2d(ctx->upic, ctx->u, n * sizeof(double), S(ctx)));
        if ((rc = ust_assemble(ctx, gravity))) return rc;
        if ((rc = ust_solve(ctx, opts, last))) return rc;
        /* the first solve enables yielding (SystemLinearEquations.c:808-809) */
        ctx->prev_solution = 1;
        if (it > 0) {
            double dd, uu;
            CKD(ctx, ust_k_sub_f64(ctx->u, ctx->upic, ctx->vt, n, S(ctx)));
            CKD(ctx, dotv(ctx, ctx->vt, ctx->vt, n, SC_T0, &dd));
            CKD(ctx, dotv(ctx, ctx->u, ctx->u, n, SC_T0, &uu));
            change = uu > 0 ? sqrt(dd / uu) : 0.0; /* SystemLinearEquations.c:852-855 */
            if (opts && opts->verbose) fprintf(stderr, "[ust] picard %d change %.3e\n", it, change);
            if (change < tol) {
                it++;
                break;
            }
        }
    }
    if (its_done) *its_done = it;
    if (final_change) *final_change = change;
    return UST_OK;
}
