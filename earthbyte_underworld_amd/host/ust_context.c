/*
 * ust_context.c -- C host layer: context, mesh, materials, swarm, fields, BCs, assembly
 * driver and the operator entry points.  Calls only the thin kernel ABI (ust_kernels.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ust_internal.h"

static void dfree(void **p) {
    if (*p) ust_dev_free(*p);
    *p = NULL;
}
#define DFREE(x) dfree((void **)&(x))

int ust_dev_alloc_zero(ust_context *ctx, void **p, size_t bytes) {
    /* round up to whole doubles so that the zero fill stays inside the allocation */
    const size_t ab = (bytes + 7) & ~(size_t)7;
    int rc = ust_dev_malloc(p, ab);
    if (rc != UST_OK) UST_FAIL(ctx, rc, "device allocation of %zu bytes failed: %s", ab, ust_dev_error_string());
    CKD(ctx, ust_k_fill_f64((double *)*p, 0.0, ab / 8, ctx->stream));
    return UST_OK;
}
#define DALLOC(ctx, ptr, bytes) CKD(ctx, ust_dev_alloc_zero(ctx, (void **)&(ptr), (bytes)))

void ust_solve_opts_default(ust_solve_opts *o) {
    memset(o, 0, sizeof(*o));
    o->precision = UST_PREC_FP64;
    o->schur_pc = UST_SCHUR_UW;
    o->mg_levels = 0;
    o->smoother_degree = 2;
    o->cheby_ratio = 12.0;
    o->outer_rtol = 1e-7;
    o->inner_rtol = 1e-9;
    o->max_outer = 1000;
    o->max_inner = 500;
    o->remove_nullspace = -1;
    o->stall_window = 30;
    o->inner_relax_cap = 1e-2;
    o->verbose = 0;
    o->vcycle_bf16 = 1;
}

int ust_create(ust_context **out, int device) {
    *out = NULL;
    ust_context *ctx = (ust_context *)calloc(1, sizeof(ust_context));
    if (!ctx) return UST_ENOMEM;
    ctx->device = device;
    ctx->buoy.vertical_axis = 1;
    /* UST_NO_GRAPH=1 launches the inner Krylov iteration kernel by kernel instead of as a captured graph
     * (same kernels and arguments; for tools that do not follow graph launches) */
    const char *ng = getenv("UST_NO_GRAPH");
    ctx->use_graphs = !(ng && ng[0] == '1');
    int rc = ust_dev_set(device);
    if (rc == UST_OK) rc = ust_stream_create(&ctx->stream);
    if (rc != UST_OK) {
        free(ctx);
        return rc;
    }
    rc = ust_dev_malloc((void **)&ctx->partial, 1024 * sizeof(double));
    if (rc == UST_OK) rc = ust_dev_malloc((void **)&ctx->scal, SC_COUNT * sizeof(double));
    if (rc == UST_OK) rc = ust_host_malloc_pinned((void **)&ctx->scal_h, SC_COUNT * sizeof(double));
    if (rc == UST_OK) rc = ust_dev_malloc((void **)&ctx->mats_d, UST_MAX_MATERIALS * sizeof(ust_material));
    if (rc != UST_OK) {
        ust_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return UST_OK;
}

const char *ust_error(const ust_context *ctx) { return ctx ? ctx->err : "null context"; }
void *ust_get_stream(ust_context *ctx) { return ctx->stream; }

static void free_mesh_data(ust_context *ctx) {
    DFREE(ctx->T);
    DFREE(ctx->u);
    DFREE(ctx->p);
    DFREE(ctx->mask_d);
    DFREE(ctx->bcval_d);
    free(ctx->mask_h);
    free(ctx->bcval_h);
    ctx->mask_h = NULL;
    ctx->bcval_h = NULL;
    DFREE(ctx->WK);
    DFREE(ctx->WG);
    DFREE(ctx->WF);
    DFREE(ctx->approxS);
    DFREE(ctx->f);
    DFREE(ctx->h);
    DFREE(ctx->tmpv);
    DFREE(ctx->tmpp);
    DFREE(ctx->kwork);
    DFREE(ctx->WK32_0);
    ctx->wk32_0_valid = 0;
    ust_graphs_reset(ctx);
    DFREE(ctx->vr); DFREE(ctx->vz); DFREE(ctx->vp); DFREE(ctx->vq); DFREE(ctx->vt); DFREE(ctx->vx);
    DFREE(ctx->va); DFREE(ctx->vb);
    DFREE(ctx->wr); DFREE(ctx->wz); DFREE(ctx->wp); DFREE(ctx->wq); DFREE(ctx->wx);
    DFREE(ctx->pr); DFREE(ctx->pz); DFREE(ctx->pd); DFREE(ctx->pq); DFREE(ctx->ph); DFREE(ctx->ppc);
    DFREE(ctx->pprev); DFREE(ctx->uprev); DFREE(ctx->upic);
    DFREE(ctx->feat); DFREE(ctx->mlp_out); DFREE(ctx->act1); DFREE(ctx->act2); DFREE(ctx->user_scale);
    ctx->feat_cap = 0;
    free(ctx->ke_cache); free(ctx->ge_cache); free(ctx->fe_cache); free(ctx->as_cache);
    ctx->ke_cache = ctx->ge_cache = ctx->fe_cache = ctx->as_cache = NULL;
    ctx->cache_valid = 0;
    ust_levels_free(ctx);
    ctx->has_bcs = 0;
    ctx->assembled = 0;
}

static void free_integration_points(ust_context *ctx) {
    DFREE(ctx->ixi0); DFREE(ctx->ixi1); DFREE(ctx->ixi2);
}

static void free_swarm(ust_context *ctx) {
    DFREE(ctx->cell_off);
    DFREE(ctx->xi0); DFREE(ctx->xi1); DFREE(ctx->xi2); DFREE(ctx->w);
    free_integration_points(ctx);
    DFREE(ctx->mat); DFREE(ctx->coh); DFREE(ctx->eta); DFREE(ctx->yielded);
    DFREE(ctx->tensile); DFREE(ctx->pfw); DFREE(ctx->pfw_inc);
    ctx->np = 0;
}

void ust_destroy(ust_context *ctx) {
    if (!ctx) return;
    if (ctx->stream) ust_stream_sync(ctx->stream);
    free_mesh_data(ctx);
    free_swarm(ctx);
    ust_mlp_free(ctx);
    DFREE(ctx->partial);
    DFREE(ctx->scal);
    DFREE(ctx->mats_d);
    if (ctx->scal_h) ust_host_free_pinned(ctx->scal_h);
    if (ctx->stream) ust_stream_destroy(ctx->stream);
    free(ctx);
}

int ust_read_scalar(ust_context *ctx, int slot, double *v) {
    CKD(ctx, ust_memcpy_d2h(ctx->scal_h + slot, ctx->scal + slot, sizeof(double), ctx->stream));
    *v = ctx->scal_h[slot];
    return UST_OK;
}

int ust_set_mesh(ust_context *ctx, int dim, const int nel[3], const double xmin[3], const double xmax[3]) {
    if (dim != 2 && dim != 3) UST_FAIL(ctx, UST_EINVAL, "dim must be 2 or 3 (got %d)", dim);
    for (int d = 0; d < dim; d++) {
        if (nel[d] < 1) UST_FAIL(ctx, UST_EINVAL, "elementRes[%d] must be >= 1", d);
        if (!(xmax[d] > xmin[d])) UST_FAIL(ctx, UST_EINVAL, "max <= min on axis %d", d);
    }
    free_mesh_data(ctx);
    memset(&ctx->grid, 0, sizeof(ctx->grid));
    ctx->grid.dim = dim;
    ctx->nnodes = 1;
    ctx->nel = 1;
    for (int d = 0; d < 3; d++) {
        ctx->grid.nel[d] = d < dim ? nel[d] : 1;
        ctx->grid.xmin[d] = d < dim ? xmin[d] : 0.0;
        ctx->grid.h[d] = d < dim ? (xmax[d] - xmin[d]) / nel[d] : 1.0;
        if (d < dim) {
            ctx->nnodes *= nel[d] + 1;
            ctx->nel *= nel[d];
        }
    }
    ctx->grid.ymax = xmax[1];
    for (int d = 0; d < 3; d++) ctx->grid.xmax[d] = d < dim ? xmax[d] : 1.0;
    ctx->ndof = ctx->nnodes * dim;
    size_t nd = (size_t)ctx->ndof, ne = (size_t)ctx->nel;
    DALLOC(ctx, ctx->T, (size_t)ctx->nnodes * sizeof(double));
    DALLOC(ctx, ctx->u, nd * sizeof(double));
    DALLOC(ctx, ctx->p, ne * sizeof(double));
    DALLOC(ctx, ctx->mask_d, nd);
    DALLOC(ctx, ctx->bcval_d, nd * sizeof(double));
    ctx->mask_h = (unsigned char *)calloc(nd, 1);
    ctx->bcval_h = (double *)calloc(nd, sizeof(double));
    int nq = dim == 3 ? 27 : 9, nv = 1 << dim;
    (void)nq;
    DALLOC(ctx, ctx->WK, ust_k_weights_len(&ctx->grid) * sizeof(double));
    DALLOC(ctx, ctx->WG, (size_t)nv * ne * sizeof(double));
    DALLOC(ctx, ctx->WF, (size_t)nv * ne * sizeof(double));
    DALLOC(ctx, ctx->approxS, ne * sizeof(double));
    DALLOC(ctx, ctx->f, nd * sizeof(double));
    DALLOC(ctx, ctx->h, ne * sizeof(double));
    DALLOC(ctx, ctx->tmpv, nd * sizeof(double));
    DALLOC(ctx, ctx->tmpp, ne * sizeof(double));
    ctx->kwork_len = ust_k_apply_work_len(&ctx->grid);
    DALLOC(ctx, ctx->kwork, ctx->kwork_len * sizeof(double));
    ctx->has_mesh = 1;
    ctx->prev_solution = 0;
    ctx->nullspace = 0;
    return UST_OK;
}

int ust_set_materials(ust_context *ctx, int nmat, const ust_material *mats) {
    if (nmat < 1 || nmat > UST_MAX_MATERIALS) UST_FAIL(ctx, UST_EINVAL, "nmat must be in [1,%d]", UST_MAX_MATERIALS);
    for (int m = 0; m < nmat; m++) {
        if (mats[m].nrh < 1 || mats[m].nrh > UST_MAX_RHEOLOGIES)
            UST_FAIL(ctx, UST_EINVAL, "material %d: no rheology (\"No material available\", ConstitutiveMatrix.c:301)", m);
        for (int r = 0; r < mats[m].nrh; r++)
            if (mats[m].rh[r].type < UST_RH_MATERIALVISCOSITY || mats[m].rh[r].type > UST_RH_DRUCKERPRAGER)
                UST_FAIL(ctx, UST_EINVAL, "material %d rheology %d: unknown type %d", m, r, mats[m].rh[r].type);
    }
    memcpy(ctx->mats_h, mats, sizeof(ust_material) * nmat);
    ctx->nmat = nmat;
    CKD(ctx, ust_memcpy_h2d(ctx->mats_d, mats, sizeof(ust_material) * nmat, ctx->stream));
    ctx->assembled = 0;
    return UST_OK;
}

static int streq(const char *a, const char *b) { return strcmp(a, b) == 0; }

int ust_rheology_from_xml(const char *type_name, int nkeys, const char *const *keys, const double *values,
                          ust_rheology *out) {
    memset(out, 0, sizeof(*out));
    /* defaults as the reference's Stg_ComponentFactory_GetDouble calls */
    if (streq(type_name, "MaterialViscosity")) {
        out->type = UST_RH_MATERIALVISCOSITY;
        out->p[0] = 1.0; /* eta0 */
    } else if (streq(type_name, "Arrhenius")) {
        out->type = UST_RH_ARRHENIUS;
        out->p[0] = 1.0; /* eta0 */
        out->p[1] = 0.0; /* activationEnergy */
        out->p[2] = 0.0; /* activationVolume */
        out->p[3] = 1.0; /* referenceTemperature */
    } else if (streq(type_name, "ViscosityLimiter")) {
        out->type = UST_RH_VISCOSITYLIMITER;
        out->p[0] = 1e4;  /* maxViscosity */
        out->p[1] = 1e-2; /* minViscosity */
    } else if (streq(type_name, "VonMises")) {
        out->type = UST_RH_VONMISES;
    } else if (streq(type_name, "DruckerPrager")) {
        out->type = UST_RH_DRUCKERPRAGER;
    } else {
        return UST_EINVAL;
    }
    for (int k = 0; k < nkeys; k++) {
        const char *key = keys[k];
        double v = values[k];
        int t = out->type, slot = -1;
        if (t == UST_RH_MATERIALVISCOSITY && streq(key, "eta0")) slot = 0;
        if (t == UST_RH_ARRHENIUS) {
            if (streq(key, "eta0")) slot = 0;
            if (streq(key, "activationEnergy")) slot = 1;
            if (streq(key, "activationVolume")) slot = 2;
            if (streq(key, "referenceTemperature")) slot = 3;
        }
        if (t == UST_RH_VISCOSITYLIMITER) {
            if (streq(key, "maxViscosity")) slot = 0;
            if (streq(key, "minViscosity")) slot = 1;
        }
        if (t == UST_RH_VONMISES) {
            if (streq(key, "cohesion")) slot = 0;
            if (streq(key, "minimumViscosity")) slot = 1;
            if (streq(key, "strainRateSoftening")) slot = 2;
            if (streq(key, "cohesionAfterSoftening")) slot = 3;
            if (streq(key, "StrainWeakening")) slot = 4;
        }
        if (t == UST_RH_DRUCKERPRAGER) {
            if (streq(key, "cohesion")) slot = 0;
            if (streq(key, "frictionCoefficient")) slot = 1;
            if (streq(key, "minimumYieldStress")) slot = 2;
            if (streq(key, "minimumViscosity")) slot = 3;
            if (streq(key, "strainRateSoftening")) slot = 4;
            if (streq(key, "cohesionAfterSoftening")) slot = 5;
            if (streq(key, "frictionCoefficientAfterSoftening")) slot = 6;
            if (streq(key, "StrainWeakening")) slot = 7;
        }
        if (slot < 0) return UST_EINVAL;
        out->p[slot] = v;
    }
    return UST_OK;
}

int ust_swarm_upload_dev(ust_context *ctx, int np, const int *cell_off_dev, const double *xi0_dev,
                         const double *xi1_dev, const double *xi2_dev, const double *w_dev,
                         const unsigned char *material_dev, const float *coh_scale_dev, int max_ppc) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "ust_set_mesh must come before the swarm");
    if (np < 0 || !cell_off_dev || !xi0_dev || !xi1_dev || !w_dev || (ctx->grid.dim == 3 && !xi2_dev))
        UST_FAIL(ctx, UST_EINVAL, "swarm arrays missing");
    free_swarm(ctx);
    size_t n = (size_t)np;
    ctx->np = np;
    ctx->max_ppc = max_ppc;
    DALLOC(ctx, ctx->cell_off, (size_t)(ctx->nel + 1) * sizeof(int));
    DALLOC(ctx, ctx->xi0, n * sizeof(double));
    DALLOC(ctx, ctx->xi1, n * sizeof(double));
    if (ctx->grid.dim == 3) DALLOC(ctx, ctx->xi2, n * sizeof(double));
    DALLOC(ctx, ctx->w, n * sizeof(double));
    DALLOC(ctx, ctx->eta, n * sizeof(double));
    DALLOC(ctx, ctx->yielded, n);
    DALLOC(ctx, ctx->tensile, n);
    CKD(ctx, ust_memcpy_d2d(ctx->cell_off, cell_off_dev, (size_t)(ctx->nel + 1) * sizeof(int), ctx->stream));
    CKD(ctx, ust_memcpy_d2d(ctx->xi0, xi0_dev, n * sizeof(double), ctx->stream));
    CKD(ctx, ust_memcpy_d2d(ctx->xi1, xi1_dev, n * sizeof(double), ctx->stream));
    if (ctx->grid.dim == 3) CKD(ctx, ust_memcpy_d2d(ctx->xi2, xi2_dev, n * sizeof(double), ctx->stream));
    CKD(ctx, ust_memcpy_d2d(ctx->w, w_dev, n * sizeof(double), ctx->stream));
    if (material_dev) {
        DALLOC(ctx, ctx->mat, n);
        CKD(ctx, ust_memcpy_d2d(ctx->mat, material_dev, n, ctx->stream));
    }
    if (coh_scale_dev) {
        DALLOC(ctx, ctx->coh, n * sizeof(float));
        CKD(ctx, ust_memcpy_d2d(ctx->coh, coh_scale_dev, n * sizeof(float), ctx->stream));
    }
    CKD(ctx, ust_stream_sync(ctx->stream));
    ctx->assembled = 0;
    return UST_OK;
}

int ust_swarm_upload(ust_context *ctx, int np, const int *cell_off, const double *xi, const double *w,
                     const unsigned char *material, const float *coh_scale) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "ust_set_mesh must come before the swarm");
    if (np < 0 || !cell_off || !xi || !w) UST_FAIL(ctx, UST_EINVAL, "swarm arrays missing");
    if (cell_off[0] != 0 || cell_off[ctx->nel] != np) UST_FAIL(ctx, UST_EINVAL, "cell_off must run 0..np");
    int max_ppc = 0;
    for (int e = 0; e < ctx->nel; e++) {
        int c = cell_off[e + 1] - cell_off[e];
        if (c < 0) UST_FAIL(ctx, UST_EINVAL, "cell_off not monotone at element %d", e);
        if (c == 0)
            UST_FAIL(ctx, UST_EINVAL, "element %d has no particles (ConstantWeights divides by the cell count)", e);
        if (c > max_ppc) max_ppc = c;
    }
    if (material)
        for (int i = 0; i < np; i++)
            if (material[i] >= ctx->nmat)
                UST_FAIL(ctx, UST_EINVAL, "particle %d: material %d not defined (\"No material available\")", i,
                         material[i]);
    size_t n = (size_t)np;
    double *soa = (double *)malloc(n * sizeof(double) * 3 + 8);
    if (!soa) UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    for (size_t i = 0; i < n; i++) {
        soa[i] = xi[3 * i];
        soa[n + i] = xi[3 * i + 1];
        soa[2 * n + i] = xi[3 * i + 2];
    }
    void *d_off = NULL, *d_xi = NULL, *d_w = NULL, *d_m = NULL, *d_c = NULL;
    int rc = UST_OK;
    rc = ust_dev_malloc(&d_off, (size_t)(ctx->nel + 1) * sizeof(int));
    if (!rc) rc = ust_dev_malloc(&d_xi, 3 * n * sizeof(double) + 8);
    if (!rc) rc = ust_dev_malloc(&d_w, n * sizeof(double) + 8);
    if (!rc && material) rc = ust_dev_malloc(&d_m, n + 8);
    if (!rc && coh_scale) rc = ust_dev_malloc(&d_c, n * sizeof(float) + 8);
    if (!rc) rc = ust_memcpy_h2d(d_off, cell_off, (size_t)(ctx->nel + 1) * sizeof(int), ctx->stream);
    if (!rc) rc = ust_memcpy_h2d(d_xi, soa, 3 * n * sizeof(double), ctx->stream);
    if (!rc) rc = ust_memcpy_h2d(d_w, w, n * sizeof(double), ctx->stream);
    if (!rc && material) rc = ust_memcpy_h2d(d_m, material, n, ctx->stream);
    if (!rc && coh_scale) rc = ust_memcpy_h2d(d_c, coh_scale, n * sizeof(float), ctx->stream);
    free(soa);
    if (!rc)
        rc = ust_swarm_upload_dev(ctx, np, (const int *)d_off, (const double *)d_xi, (const double *)d_xi + n,
                                  (const double *)d_xi + 2 * n, (const double *)d_w, (const unsigned char *)d_m,
                                  (const float *)d_c, max_ppc);
    else
        snprintf(ctx->err, sizeof(ctx->err), "swarm upload failed: %s", ust_dev_error_string());
    ust_dev_free(d_off);
    ust_dev_free(d_xi);
    ust_dev_free(d_w);
    ust_dev_free(d_m);
    ust_dev_free(d_c);
    return rc;
}

void ust_buoyancy_direction(const ust_context *ctx, double dir[3]) {
    for (int d = 0; d < 3; d++) dir[d] = 0.0;
    if (ctx->buoy.use_direction) {
        for (int d = 0; d < 3; d++) dir[d] = ctx->buoy.direction[d];
    } else {
        dir[ctx->buoy.vertical_axis] = -1.0;
    }
}

int ust_set_buoyancy(ust_context *ctx, const ust_buoyancy *b) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "no mesh");
    if (!b) {
        memset(&ctx->buoy, 0, sizeof(ctx->buoy));
        ctx->buoy.vertical_axis = 1;
    } else {
        if (b->vertical_axis < 0 || b->vertical_axis >= ctx->grid.dim)
            UST_FAIL(ctx, UST_EINVAL, "VerticalAxis %d outside 0..%d", b->vertical_axis, ctx->grid.dim - 1);
        if (b->hydrostatic && b->hs_v != 0 && !(b->hs_lower_density != 0.0 && b->hs_width + b->hs_v * b->hs_time != 0.0))
            UST_FAIL(ctx, UST_EINVAL, "HydrostaticTerm: a moving boundary needs a lower density and width + v t != 0");
        ctx->buoy = *b;
    }
    ctx->assembled = 0;
    return UST_OK;
}

/* ---- StrainWeakening ---- */
void ust_strain_weakening_default(ust_strain_weakening *sw) {
    memset(sw, 0, sizeof(*sw));
    sw->softening_strain = HUGE_VAL;
    sw->initial_damage_wavenumber = -1.0;
    for (int d = 0; d < 3; d++) sw->wavenumber_sin[d] = sw->wavenumber_cos[d] = -1.0;
    sw->initial_damage_factor = 1.0;
}

/* the per-particle state arrays, zero when first allocated for the current swarm */
static int sw_state_alloc(ust_context *ctx) {
    const size_t n = (size_t)ctx->np;
    if (!ctx->pfw) DALLOC(ctx, ctx->pfw, n * sizeof(double) + 8);
    if (!ctx->pfw_inc) DALLOC(ctx, ctx->pfw_inc, n * sizeof(double) + 8);
    return UST_OK;
}

int ust_set_strain_weakening(ust_context *ctx, const ust_strain_weakening *sw) {
    if (!sw) {
        ctx->sw_on = 0;
        DFREE(ctx->pfw);
        DFREE(ctx->pfw_inc);
        ctx->assembled = 0;
        return UST_OK;
    }
    if (!(sw->softening_strain > 0.0)) UST_FAIL(ctx, UST_EINVAL, "softeningStrain must be > 0");
    ctx->sw = *sw;
    ctx->sw_on = 1;
    if (ctx->np) {
        int rc = sw_state_alloc(ctx);
        if (rc) return rc;
    }
    ctx->assembled = 0;
    return UST_OK;
}

static int sw_shape_inside(const ust_shape *sh, const double *x, int dim) {
    if (sh->type == UST_SHAPE_EVERYWHERE) return 1;
    if (sh->type == UST_SHAPE_BELOWCOSINEPLANE)
        return x[1] <= sh->p[0] + sh->p[1] * cos((2 * M_PI * x[0] / sh->p[2]) + sh->p[3]);
    if (sh->type == UST_SHAPE_SPHERE) {
        double r2 = 0.0;
        for (int d = 0; d < dim; d++) r2 += (x[d] - sh->p[d]) * (x[d] - sh->p[d]);
        return r2 <= sh->p[3] * sh->p[3];
    }
    if (sh->type == UST_SHAPE_BOX) {
        for (int d = 0; d < dim; d++)
            if (!(x[d] >= sh->p[d] && x[d] <= sh->p[3 + d])) return 0;
        return 1;
    }
    return 0;
}

int ust_strain_weakening_initialise(ust_context *ctx) {
    if (!ctx->sw_on) UST_FAIL(ctx, UST_ESTATE, "strain weakening is off (ust_set_strain_weakening)");
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    const int dim = ctx->grid.dim, nel = ctx->nel, np = ctx->np;
    const ust_strain_weakening *sw = &ctx->sw;
    int *off = (int *)malloc((size_t)(nel + 1) * sizeof(int));
    double *xi = (double *)malloc((size_t)np * 3 * sizeof(double));
    double *pfw = (double *)malloc((size_t)np * sizeof(double));
    if (!off || !xi || !pfw) {
        free(off); free(xi); free(pfw);
        UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    }
    int rc = ust_swarm_download(ctx, off, xi, NULL, NULL);
    if (rc) {
        free(off); free(xi); free(pfw);
        return rc;
    }
    /* the material points in swarm order; srand(randomSeed) either way (StrainWeakening.c:338-345) */
    srand((unsigned)sw->random_seed);
    for (int e = 0; e < nel; e++) {
        int idx[3] = {e % ctx->grid.nel[0], (e / ctx->grid.nel[0]) % ctx->grid.nel[1],
                      dim == 3 ? e / (ctx->grid.nel[0] * ctx->grid.nel[1]) : 0};
        for (int p = off[e]; p < off[e + 1]; p++) {
            double v = 0.0;
            if (rand() < (RAND_MAX * sw->initial_damage_fraction)) {
                double x[3] = {0.0, 0.0, 0.0};
                for (int d = 0; d < dim; d++) {
                    const double lo = ctx->grid.xmin[d] + idx[d] * ctx->grid.h[d];
                    x[d] = lo + 0.5 * (xi[3 * (size_t)p + d] + 1.0) * ctx->grid.h[d];
                }
                if (sw->has_shape && !sw_shape_inside(&sw->initial_strain_shape, x, dim)) {
                    pfw[p] = 0.0;
                    continue;
                }
                v = sw->initial_damage_factor * rand() * sw->softening_strain / RAND_MAX;
                if (sw->initial_damage_wavenumber > 0.0 && sw->wavenumber_cos[0] == -1.0)
                    v *= pow(cos(M_PI * x[0] * sw->initial_damage_wavenumber), 2.0);
                /* order of the reference: CosI, SinI, CosJ, SinJ, CosK, SinK */
                for (int d = 0; d < 3; d++) {
                    if (sw->wavenumber_cos[d] > 0.0) v *= pow(cos(M_PI * x[d] * sw->wavenumber_cos[d]), 2.0);
                    if (sw->wavenumber_sin[d] > 0.0) v *= pow(sin(M_PI * x[d] * sw->wavenumber_sin[d]), 2.0);
                }
            }
            pfw[p] = v;
        }
    }
    rc = sw_state_alloc(ctx);
    if (!rc) rc = ust_memcpy_h2d(ctx->pfw, pfw, (size_t)np * sizeof(double), ctx->stream);
    if (!rc) rc = ust_k_fill_f64(ctx->pfw_inc, 0.0, (size_t)np, ctx->stream);
    if (!rc) rc = ust_stream_sync(ctx->stream);
    free(off); free(xi); free(pfw);
    if (rc) UST_FAIL(ctx, rc, "strain weakening set-up failed: %s", ust_dev_error_string());
    ctx->assembled = 0;
    return UST_OK;
}

int ust_strain_weakening_step(ust_context *ctx, double dt) {
    if (!ctx->sw_on || !ctx->pfw) UST_FAIL(ctx, UST_ESTATE, "strain weakening is off");
    CKD(ctx, ust_k_strain_weakening_step(ctx->pfw, ctx->pfw_inc, ctx->np, dt, ctx->stream));
    ctx->assembled = 0;
    return UST_OK;
}

int ust_get_strain_weakening(ust_context *ctx, double *pfw, double *pfw_inc) {
    if (!ctx->sw_on || !ctx->pfw) UST_FAIL(ctx, UST_ESTATE, "strain weakening is off");
    const size_t n = (size_t)ctx->np * sizeof(double);
    if (pfw) CKD(ctx, ust_memcpy_d2h(pfw, ctx->pfw, n, ctx->stream));
    if (pfw_inc) CKD(ctx, ust_memcpy_d2h(pfw_inc, ctx->pfw_inc, n, ctx->stream));
    return UST_OK;
}

int ust_set_strain_weakening_state(ust_context *ctx, const double *pfw, const double *pfw_inc) {
    if (!ctx->sw_on) UST_FAIL(ctx, UST_ESTATE, "strain weakening is off");
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    int rc = sw_state_alloc(ctx);
    if (rc) return rc;
    const size_t n = (size_t)ctx->np * sizeof(double);
    if (pfw) CKD(ctx, ust_memcpy_h2d(ctx->pfw, pfw, n, ctx->stream));
    if (pfw_inc) CKD(ctx, ust_memcpy_h2d(ctx->pfw_inc, pfw_inc, n, ctx->stream));
    CKD(ctx, ust_stream_sync(ctx->stream));
    ctx->assembled = 0;
    return UST_OK;
}

int ust_get_tensile_failure(ust_context *ctx, unsigned char *tensile) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    CKD(ctx, ust_memcpy_d2h(tensile, ctx->tensile, (size_t)ctx->np, ctx->stream));
    return UST_OK;
}

int ust_set_temperature(ust_context *ctx, const double *T) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "no mesh");
    CKD(ctx, ust_memcpy_h2d(ctx->T, T, (size_t)ctx->nnodes * sizeof(double), ctx->stream));
    ctx->assembled = 0;
    return UST_OK;
}

int ust_set_solution(ust_context *ctx, const double *u, const double *p) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "no mesh");
    if (u) CKD(ctx, ust_memcpy_h2d(ctx->u, u, (size_t)ctx->ndof * sizeof(double), ctx->stream));
    if (p) CKD(ctx, ust_memcpy_h2d(ctx->p, p, (size_t)ctx->nel * sizeof(double), ctx->stream));
    return UST_OK;
}

int ust_set_previous_solution_exists(ust_context *ctx, int flag) {
    ctx->prev_solution = flag ? 1 : 0;
    return UST_OK;
}

int ust_set_bcs(ust_context *ctx, const unsigned char *mask, const double *values) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "no mesh");
    memcpy(ctx->mask_h, mask, (size_t)ctx->ndof);
    ctx->bc_nonzero = 0;
    for (int i = 0; i < ctx->ndof; i++) {
        ctx->bcval_h[i] = (mask[i] && values) ? values[i] : 0.0;
        if (ctx->bcval_h[i] != 0.0) ctx->bc_nonzero = 1;
    }
    CKD(ctx, ust_memcpy_h2d(ctx->mask_d, ctx->mask_h, (size_t)ctx->ndof, ctx->stream));
    CKD(ctx, ust_memcpy_h2d(ctx->bcval_d, ctx->bcval_h, (size_t)ctx->ndof * sizeof(double), ctx->stream));
    ctx->has_bcs = 1;
    ust_levels_free(ctx);
    ctx->assembled = 0;
    return UST_OK;
}

/* K, G, approxS and RHS assembly.  The particle kernel runs the rheology chain of each
 * particle and reduces the quadrature weights; everything else derives from the weights. */
int ust_assemble(ust_context *ctx, double gravity) {
    if (!ctx->has_mesh || !ctx->np || !ctx->nmat) UST_FAIL(ctx, UST_ESTATE, "mesh, materials and swarm are required");
    if (!ctx->has_bcs) {
        unsigned char *m0 = (unsigned char *)calloc((size_t)ctx->ndof, 1);
        int rc = ust_set_bcs(ctx, m0, NULL);
        free(m0);
        if (rc) return rc;
    }
    size_t nd = (size_t)ctx->ndof, ne = (size_t)ctx->nel;
    /* the temperature and strain-rate/pressure lookups are compiled into the particle kernel only when a
     * material's rheology chain or density reads them; the density law runs only when a material has a
     * thermal or depth coefficient or a hydrostatic background is set (else rho = the material density) */
    int needT = 0, needV = 0, fullDensity = ctx->buoy.hydrostatic != 0;
    for (int m = 0; m < ctx->nmat; m++) {
        for (int r = 0; r < ctx->mats_h[m].nrh; r++) {
            int t = ctx->mats_h[m].rh[r].type;
            if (t == UST_RH_ARRHENIUS) needT = 1;
            if (t == UST_RH_VONMISES || t == UST_RH_DRUCKERPRAGER) needV = 1;
        }
        if (ctx->mats_h[m].alpha != 0.0) needT = fullDensity = 1;
        if (ctx->mats_h[m].beta != 0.0) fullDensity = 1;
    }
    const int ppc_mean = (ctx->np + ctx->nel - 1) / ctx->nel;
    ust_yield_state ys = {NULL, NULL, ctx->tensile, 1.0, 0.0};
    if (ctx->sw_on) {
        int rc = sw_state_alloc(ctx);
        if (rc) return rc;
        ys.pfw = ctx->pfw;
        ys.pfw_inc = ctx->pfw_inc;
        ys.softening_strain = ctx->sw.softening_strain;
        ys.healing_rate = ctx->sw.healing_rate;
    }
    /* the integration points: the material points themselves, or the points a weights calculator moved */
    const int sep = ctx->ixi0 != NULL;
    const double *q0 = sep ? ctx->ixi0 : ctx->xi0, *q1 = sep ? ctx->ixi1 : ctx->xi1, *q2 = sep ? ctx->ixi2 : ctx->xi2;
    CKD(ctx, ust_k_particle_weights(&ctx->grid, ctx->cell_off, ppc_mean, q0, q1, q2, ctx->w,
                                    ctx->mat, ctx->coh, ctx->mats_d, ctx->nmat, needT ? ctx->T : NULL,
                                    needV ? ctx->u : NULL, needV ? ctx->p : NULL, ctx->prev_solution, &ys,
                                    fullDensity ? &ctx->buoy : NULL, ctx->eta, ctx->yielded, ctx->WK, ctx->WG,
                                    ctx->WF, ctx->stream));
    if (ctx->buoy.adjust) CKD(ctx, ust_k_adjust_force_weights(&ctx->grid, ctx->WG, ctx->WF, ctx->stream));
    CKD(ctx, ust_k_uzawa_diag(&ctx->grid, ctx->WK, ctx->WG, ctx->approxS, ctx->stream));
    /* f = buoyancy - K_{free,bc} u_bc ; h = -G_bc^T u_bc */
    CKD(ctx, ust_k_fill_f64(ctx->f, 0.0, nd, ctx->stream));
    double dir[3];
    ust_buoyancy_direction(ctx, dir);
    CKD(ctx, ust_k_force_dir(&ctx->grid, ctx->WF, gravity, dir, ctx->mask_d, ctx->f, ctx->stream));
    CKD(ctx, ust_k_fill_f64(ctx->h, 0.0, ne, ctx->stream));
    if (ctx->bc_nonzero) { /* homogeneous Dirichlet values lift nothing */
        CKD(ctx, ust_k_lift_K(&ctx->grid, ctx->WK, ctx->mask_d, ctx->bcval_d, ctx->f, ctx->stream));
        CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, ctx->bcval_d, ctx->h, 1, ctx->stream));
    }
    /* is the constant pressure t in the null space of G?  ||G t|| / ||t|| < nstol ||D||_inf, nstol = 1e-7
     * (KSPRemovePressureNullspace_BSSCR, ksp_pressure_nullspace.c:139-158; BSSCR.c:391).  With particle
     * quadrature G t is usually not zero, and the constant is then part of the solution. */
    {
        double g1, gnorm;
        CKD(ctx, ust_k_fill_f64(ctx->tmpp, 1.0, ne, ctx->stream));
        CKD(ctx, ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, ctx->tmpp, ctx->tmpv, ctx->stream));
        CKD(ctx, ust_k_dot_f64(ctx->tmpv, ctx->tmpv, nd, ctx->partial, ctx->scal + SC_T0, ctx->stream));
        CKD(ctx, ust_read_scalar(ctx, SC_T0, &g1));
        CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, ctx->tmpv, ctx->tmpp, 2, ctx->stream));
        CKD(ctx, ust_k_absmax_f64(ctx->tmpp, ne, ctx->partial, ctx->scal + SC_T0, ctx->stream));
        CKD(ctx, ust_read_scalar(ctx, SC_T0, &gnorm));
        ctx->ns_ratio = gnorm > 0 ? sqrt(g1) / sqrt((double)ne) / gnorm : 0.0;
        ctx->nullspace = ctx->ns_ratio < 1e-7;
    }
    ctx->gravity = gravity;
    ctx->assembled = 1;
    ctx->levels_ready = 0;
    ctx->wk32_0_valid = 0;
    ctx->cache_valid = 0;
    return UST_OK;
}

static int need_assembled(ust_context *ctx) {
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "ust_assemble has not run since the inputs changed");
    return UST_OK;
}

int ust_get_element_matrices_dev(ust_context *ctx, int first, int count, double *Ke_dev) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    if (first < 0 || count < 0 || first + count > ctx->nel) UST_FAIL(ctx, UST_EINVAL, "element range out of bounds");
    CKD(ctx, ust_k_element_matrices(&ctx->grid, ctx->WK, first, count, Ke_dev, ctx->stream));
    return UST_OK;
}

/* element batches are bounded so that the staging buffer stays modest */
#define UST_BATCH 65536

int ust_get_element_matrices(ust_context *ctx, int first, int count, double *Ke) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    if (first < 0 || count < 0 || first + count > ctx->nel) UST_FAIL(ctx, UST_EINVAL, "element range out of bounds");
    int nd = (1 << ctx->grid.dim) * ctx->grid.dim;
    double *buf = NULL;
    int bs = count < UST_BATCH ? count : UST_BATCH;
    if (bs == 0) return UST_OK;
    CKD(ctx, ust_dev_malloc((void **)&buf, (size_t)bs * nd * nd * sizeof(double)));
    for (int s = 0; s < count; s += bs) {
        int c = count - s < bs ? count - s : bs;
        rc = ust_k_element_matrices(&ctx->grid, ctx->WK, first + s, c, buf, ctx->stream);
        if (!rc) rc = ust_memcpy_d2h(Ke + (size_t)s * nd * nd, buf, (size_t)c * nd * nd * sizeof(double), ctx->stream);
        if (rc) break;
    }
    ust_dev_free(buf);
    if (rc) UST_FAIL(ctx, rc, "element matrices: %s", ust_dev_error_string());
    return UST_OK;
}

static int get_elem_vec(ust_context *ctx, int first, int count, double *out, int which) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    if (first < 0 || count < 0 || first + count > ctx->nel) UST_FAIL(ctx, UST_EINVAL, "element range out of bounds");
    if (count == 0) return UST_OK;
    int nd = (1 << ctx->grid.dim) * ctx->grid.dim;
    double *buf = NULL;
    CKD(ctx, ust_dev_malloc((void **)&buf, (size_t)count * nd * sizeof(double)));
    if (which == 0) rc = ust_k_element_G(&ctx->grid, ctx->WG, first, count, buf, ctx->stream);
    else {
        double dir[3];
        ust_buoyancy_direction(ctx, dir);
        rc = ust_k_element_f_dir(&ctx->grid, ctx->WF, ctx->gravity, dir, first, count, buf, ctx->stream);
    }
    if (!rc) rc = ust_memcpy_d2h(out, buf, (size_t)count * nd * sizeof(double), ctx->stream);
    ust_dev_free(buf);
    if (rc) UST_FAIL(ctx, rc, "element vectors: %s", ust_dev_error_string());
    return UST_OK;
}

int ust_get_element_G(ust_context *ctx, int first, int count, double *Ge) { return get_elem_vec(ctx, first, count, Ge, 0); }
int ust_get_element_f(ust_context *ctx, int first, int count, double *fe) { return get_elem_vec(ctx, first, count, fe, 1); }

int ust_get_approxS(ust_context *ctx, double *approxS) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    CKD(ctx, ust_memcpy_d2h(approxS, ctx->approxS, (size_t)ctx->nel * sizeof(double), ctx->stream));
    return UST_OK;
}

int ust_get_particle_viscosity(ust_context *ctx, double *eta, unsigned char *yielded) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    if (eta) CKD(ctx, ust_memcpy_d2h(eta, ctx->eta, (size_t)ctx->np * sizeof(double), ctx->stream));
    if (yielded) CKD(ctx, ust_memcpy_d2h(yielded, ctx->yielded, (size_t)ctx->np, ctx->stream));
    return UST_OK;
}

int ust_get_weights(ust_context *ctx, double *WK, double *WG, double *WF) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    int nq = ctx->grid.dim == 3 ? 27 : 9, nv = 1 << ctx->grid.dim;
    size_t ne = (size_t)ctx->nel;
    if (WK) {
        /* device layout is tile-blocked; hand back [q][e] */
        size_t len = ust_k_weights_len(&ctx->grid);
        double *tmp = (double *)malloc(len * sizeof(double));
        if (!tmp) UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
        rc = ust_memcpy_d2h(tmp, ctx->WK, len * sizeof(double), ctx->stream);
        if (rc) {
            free(tmp);
            UST_FAIL(ctx, rc, "weights download failed: %s", ust_dev_error_string());
        }
        for (int q = 0; q < nq; q++)
            for (int e = 0; e < ctx->nel; e++) WK[(size_t)q * ne + e] = tmp[ust_k_weights_index(&ctx->grid, e, q)];
        free(tmp);
    }
    if (WG) CKD(ctx, ust_memcpy_d2h(WG, ctx->WG, nv * ne * sizeof(double), ctx->stream));
    if (WF) CKD(ctx, ust_memcpy_d2h(WF, ctx->WF, nv * ne * sizeof(double), ctx->stream));
    return UST_OK;
}

int ust_get_device_views(ust_context *ctx, ust_grid *grid, const double **WK, const double **WG, const double **WF,
                         const unsigned char **mask) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "no mesh");
    if (grid) *grid = ctx->grid;
    if (WK) *WK = ctx->WK;
    if (WG) *WG = ctx->WG;
    if (WF) *WF = ctx->WF;
    if (mask) *mask = ctx->mask_d;
    return UST_OK;
}

int ust_get_rhs(ust_context *ctx, double *f, double *h) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    if (f) CKD(ctx, ust_memcpy_d2h(f, ctx->f, (size_t)ctx->ndof * sizeof(double), ctx->stream));
    if (h) CKD(ctx, ust_memcpy_d2h(h, ctx->h, (size_t)ctx->nel * sizeof(double), ctx->stream));
    return UST_OK;
}

int ust_apply_K_dev(ust_context *ctx, const double *u, double *y) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    CKD(ctx, ust_k_apply_K_f64(&ctx->grid, ctx->WK, ctx->mask_d, u, y, (double *)ctx->kwork, ctx->stream));
    return UST_OK;
}

int ust_apply_K_f32_dev(ust_context *ctx, const float *u, float *y) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    size_t len = ust_k_weights_len(&ctx->grid);
    if (!ctx->WK32_0) DALLOC(ctx, ctx->WK32_0, len * sizeof(float));
    if (!ctx->wk32_0_valid) {
        CKD(ctx, ust_k_copy_f64_to_f32(ctx->WK, ctx->WK32_0, len, ctx->stream));
        ctx->wk32_0_valid = 1;
    }
    CKD(ctx, ust_k_apply_K_f32(&ctx->grid, ctx->WK32_0, ctx->mask_d, u, y, (float *)ctx->kwork, ctx->stream));
    return UST_OK;
}

int ust_apply_K_f32(ust_context *ctx, const float *u, float *y) {
    size_t n = (size_t)ctx->ndof;
    float *a = NULL, *b = NULL;
    CKD(ctx, ust_dev_malloc((void **)&a, n * sizeof(float)));
    int rc = ust_dev_malloc((void **)&b, n * sizeof(float));
    if (!rc) rc = ust_memcpy_h2d(a, u, n * sizeof(float), ctx->stream);
    if (!rc) rc = ust_apply_K_f32_dev(ctx, a, b);
    if (!rc) rc = ust_memcpy_d2h(y, b, n * sizeof(float), ctx->stream);
    ust_dev_free(a);
    ust_dev_free(b);
    return rc;
}

int ust_get_cohesion_scale(ust_context *ctx, float *coh) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    if (ctx->coh) {
        CKD(ctx, ust_memcpy_d2h(coh, ctx->coh, (size_t)ctx->np * sizeof(float), ctx->stream));
    } else {
        for (int i = 0; i < ctx->np; i++) coh[i] = 1.0f;
    }
    return UST_OK;
}

static int host_op(ust_context *ctx, const double *in, size_t nin, double *out, size_t nout, int which) {
    int rc = need_assembled(ctx);
    if (rc) return rc;
    double *a = NULL, *b = NULL;
    CKD(ctx, ust_dev_malloc((void **)&a, nin * sizeof(double)));
    rc = ust_dev_malloc((void **)&b, nout * sizeof(double));
    if (!rc) rc = ust_memcpy_h2d(a, in, nin * sizeof(double), ctx->stream);
    if (!rc) {
        if (which == 0) rc = ust_k_apply_K_f64(&ctx->grid, ctx->WK, ctx->mask_d, a, b, (double *)ctx->kwork, ctx->stream);
        else if (which == 1) rc = ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, a, b, ctx->stream);
        else rc = ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, a, b, 0, ctx->stream);
    }
    if (!rc) rc = ust_memcpy_d2h(out, b, nout * sizeof(double), ctx->stream);
    ust_dev_free(a);
    ust_dev_free(b);
    if (rc) UST_FAIL(ctx, rc, "operator apply failed: %s", ust_dev_error_string());
    return UST_OK;
}

int ust_apply_K(ust_context *ctx, const double *u, double *y) {
    return host_op(ctx, u, (size_t)ctx->ndof, y, (size_t)ctx->ndof, 0);
}
int ust_apply_G(ust_context *ctx, const double *p, double *y) {
    return host_op(ctx, p, (size_t)ctx->nel, y, (size_t)ctx->ndof, 1);
}
int ust_apply_Gt(ust_context *ctx, const double *u, double *q) {
    return host_op(ctx, u, (size_t)ctx->ndof, q, (size_t)ctx->nel, 2);
}

int ust_get_solution(ust_context *ctx, double *u, double *p) {
    if (u) CKD(ctx, ust_memcpy_d2h(u, ctx->u, (size_t)ctx->ndof * sizeof(double), ctx->stream));
    if (p) CKD(ctx, ust_memcpy_d2h(p, ctx->p, (size_t)ctx->nel * sizeof(double), ctx->stream));
    return UST_OK;
}

int ust_get_solution_dev(ust_context *ctx, const double **u, const double **p) {
    if (u) *u = ctx->u;
    if (p) *p = ctx->p;
    return UST_OK;
}

/* ---- swarm generated on the device ---- */
int ust_swarm_random(ust_context *ctx, const int *cell_off_host, int ppc, long seed) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "ust_set_mesh must come before the swarm");
    int ne = ctx->nel;
    int *off = (int *)malloc(sizeof(int) * (size_t)(ne + 1));
    if (!off) UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    if (cell_off_host) {
        memcpy(off, cell_off_host, sizeof(int) * (size_t)(ne + 1));
    } else {
        if (ppc < 1) {
            free(off);
            UST_FAIL(ctx, UST_EINVAL, "particlesPerCell must be >= 1");
        }
        for (int e = 0; e <= ne; e++) off[e] = e * ppc;
    }
    int np = off[ne], max_ppc = 0;
    for (int e = 0; e < ne; e++) {
        int c = off[e + 1] - off[e];
        if (c < 1) {
            free(off);
            UST_FAIL(ctx, UST_EINVAL, "element %d has no particles", e);
        }
        if (c > max_ppc) max_ppc = c;
    }
    free_swarm(ctx);
    size_t n = (size_t)np;
    ctx->np = np;
    ctx->max_ppc = max_ppc;
    int rc = ust_dev_alloc_zero(ctx, (void **)&ctx->cell_off, (size_t)(ne + 1) * sizeof(int));
    if (!rc) rc = ust_memcpy_h2d(ctx->cell_off, off, (size_t)(ne + 1) * sizeof(int), ctx->stream);
    free(off);
    if (rc) UST_FAIL(ctx, rc, "swarm upload failed: %s", ust_dev_error_string());
    DALLOC(ctx, ctx->xi0, n * sizeof(double));
    DALLOC(ctx, ctx->xi1, n * sizeof(double));
    if (ctx->grid.dim == 3) DALLOC(ctx, ctx->xi2, n * sizeof(double));
    DALLOC(ctx, ctx->w, n * sizeof(double));
    DALLOC(ctx, ctx->eta, n * sizeof(double));
    DALLOC(ctx, ctx->yielded, n);
    DALLOC(ctx, ctx->tensile, n);
    DALLOC(ctx, ctx->mat, n);
    CKD(ctx, ust_k_random_layout(&ctx->grid, ctx->cell_off, np, seed, ctx->xi0, ctx->xi1, ctx->xi2, ctx->w,
                                 ctx->stream));
    CKD(ctx, ust_stream_sync(ctx->stream));
    ctx->assembled = 0;
    return UST_OK;
}

/* Direction numbers of the first three generators of SobolGenerator's table (SobolGenerator.c:340-470: primitive
 * polynomials x + 1, x^2 + x + 1, x^3 + x + 1 with initial numbers {1}, {1, 1}, {1, 3, 7}, Joe & Kuo's values),
 * continued by the recurrence m_i = XOR_k a_k 2^k m_{i-k} ^ 2^s m_{i-s} ^ m_{i-s} (SobolGenerator.c:245-290). */
static void sobol_direction_numbers(unsigned m[3][32]) {
    static const unsigned degree[3] = {1, 2, 3}, coefficient[3] = {0, 1, 1};
    static const unsigned init[3][3] = {{1, 0, 0}, {1, 1, 0}, {1, 3, 7}};
    for (int d = 0; d < 3; d++) {
        const unsigned s = degree[d], a = coefficient[d];
        memset(m[d], 0, sizeof(m[d]));
        for (unsigned i = 0; i < s; i++) m[d][i] = init[d][i];
        for (unsigned i = s; i < 30; i++) {
            unsigned v = 0;
            for (unsigned k = 1; k < s; k++)
                if ((a >> (s - 1 - k)) & 1u) v ^= (1u << k) * m[d][i - k];
            v ^= (1u << s) * m[d][i - s];
            v ^= m[d][i - s];
            m[d][i] = v;
        }
    }
}

int ust_swarm_spacefiller(ust_context *ctx, double avg_ppc) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "ust_set_mesh must come before the swarm");
    const int nel = ctx->nel, dim = ctx->grid.dim;
    /* GlobalParticleLayout.c:183: totalInitialParticles = cells x averageInitialParticlesPerCell, truncated */
    const double total_d = (double)(unsigned)nel * avg_ppc;
    if (!(avg_ppc > 0.0) || !(total_d >= 1.0) || total_d >= 1073741824.0)
        UST_FAIL(ctx, UST_EINVAL, "averageInitialParticlesPerCell %g gives no particles or too many", avg_ppc);
    const int np = (int)(unsigned)total_d;
    unsigned dirs[3][32];
    sobol_direction_numbers(dirs);
    const size_t n = (size_t)np, wb = ust_k_sort_by_cell_work_bytes(np, nel);
    void *key = NULL, *key_s = NULL, *perm = NULL, *perm_t = NULL, *nxi = NULL, *work = NULL, *off = NULL;
    void *oxi = NULL, *ow = NULL;
    int rc = UST_OK, max_ppc = 0;
    if (!rc) rc = ust_dev_malloc(&key, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&key_s, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&perm, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&perm_t, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&nxi, 3 * n * sizeof(double) + 8);
    if (!rc) rc = ust_dev_malloc(&work, wb + 8);
    if (!rc) rc = ust_dev_malloc(&off, (size_t)(nel + 1) * sizeof(int));
    if (!rc) rc = ust_dev_malloc(&oxi, 3 * n * sizeof(double) + 8);
    if (!rc) rc = ust_dev_malloc(&ow, n * sizeof(double) + 8);
    double *n0 = (double *)nxi, *n1 = n0 + n, *n2 = n0 + 2 * n;
    double *o0 = (double *)oxi, *o1 = o0 + n, *o2 = o0 + 2 * n;
    if (!rc) rc = ust_k_spacefiller_layout(&ctx->grid, np, &dirs[0][0], (unsigned *)key, n0, n1, n2, ctx->stream);
    if (!rc)
        rc = ust_k_sort_by_cell((const unsigned *)key, (unsigned *)key_s, (unsigned *)perm, (unsigned *)perm_t, np,
                                nel, work, wb, ctx->stream);
    if (!rc) rc = ust_k_cell_offsets((const unsigned *)key_s, np, nel, (int *)off, ctx->stream);
    int *off_h = (int *)malloc((size_t)(nel + 1) * sizeof(int));
    if (!off_h && !rc) rc = UST_ENOMEM;
    if (!rc) rc = ust_memcpy_d2h(off_h, off, (size_t)(nel + 1) * sizeof(int), ctx->stream);
    if (!rc)
        for (int e = 0; e < nel; e++)
            if (off_h[e + 1] - off_h[e] > max_ppc) max_ppc = off_h[e + 1] - off_h[e];
    free(off_h);
    if (!rc)
        rc = ust_k_swarm_gather(dim, np, (const unsigned *)perm, (const unsigned *)key_s, (const int *)off, n0, n1, n2,
                                NULL, NULL, o0, o1, o2, (double *)ow, NULL, NULL, ctx->stream);
    if (!rc) {
        rc = ust_swarm_upload_dev(ctx, np, (const int *)off, o0, o1, dim == 3 ? o2 : NULL, (const double *)ow, NULL,
                                  NULL, max_ppc);
    } else {
        snprintf(ctx->err, sizeof(ctx->err), "space-filler layout failed (%d): %s", rc, ust_dev_error_string());
    }
    void *tmp[] = {key, key_s, perm, perm_t, nxi, work, off, oxi, ow};
    for (size_t i = 0; i < sizeof(tmp) / sizeof(tmp[0]); i++) ust_dev_free(tmp[i]);
    return rc;
}

int ust_swarm_assign_materials(ust_context *ctx, int nshape, const ust_shape *shapes,
                               const unsigned char *cell_material) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    if (nshape < 0 || nshape > 4096) UST_FAIL(ctx, UST_EINVAL, "bad shape count");
    for (int s = 0; s < nshape; s++)
        if (shapes[s].material < 0 || shapes[s].material >= ctx->nmat)
            UST_FAIL(ctx, UST_EINVAL, "shape %d names material %d (only %d defined)", s, shapes[s].material, ctx->nmat);
    if (!ctx->mat) DALLOC(ctx, ctx->mat, (size_t)ctx->np);
    void *d_sh = NULL, *d_cm = NULL;
    int rc = UST_OK;
    if (nshape) {
        rc = ust_dev_malloc(&d_sh, sizeof(ust_shape) * (size_t)nshape);
        if (!rc) rc = ust_memcpy_h2d(d_sh, shapes, sizeof(ust_shape) * (size_t)nshape, ctx->stream);
    }
    if (!rc && cell_material) {
        for (int e = 0; e < ctx->nel; e++)
            if (cell_material[e] >= ctx->nmat) {
                ust_dev_free(d_sh);
                UST_FAIL(ctx, UST_EINVAL, "cell %d names material %d (only %d defined)", e, cell_material[e], ctx->nmat);
            }
        rc = ust_dev_malloc(&d_cm, (size_t)ctx->nel);
        if (!rc) rc = ust_memcpy_h2d(d_cm, cell_material, (size_t)ctx->nel, ctx->stream);
    }
    if (!rc)
        rc = ust_k_assign_materials(&ctx->grid, ctx->cell_off, ctx->np, ctx->xi0, ctx->xi1, ctx->xi2,
                                    (const ust_shape *)d_sh, nshape, (const unsigned char *)d_cm, ctx->mat, ctx->stream);
    if (!rc) rc = ust_stream_sync(ctx->stream);
    ust_dev_free(d_sh);
    ust_dev_free(d_cm);
    if (rc) UST_FAIL(ctx, rc, "material assignment failed: %s", ust_dev_error_string());
    ctx->assembled = 0;
    return UST_OK;
}

int ust_swarm_assign_materials_field(ust_context *ctx, const float *field_nodal, double lo, double hi, int nbins,
                                     int first_material) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    if (!field_nodal || nbins < 1 || !(hi > lo)) UST_FAIL(ctx, UST_EINVAL, "bad material field");
    if (first_material < 0 || first_material + nbins > ctx->nmat)
        UST_FAIL(ctx, UST_EINVAL, "materials %d..%d not defined (only %d)", first_material, first_material + nbins - 1,
                 ctx->nmat);
    if (!ctx->mat) DALLOC(ctx, ctx->mat, (size_t)ctx->np);
    void *d_f = NULL;
    const size_t nb = (size_t)ctx->nnodes * sizeof(float);
    int rc = ust_dev_malloc(&d_f, nb);
    if (!rc) rc = ust_memcpy_h2d(d_f, field_nodal, nb, ctx->stream);
    if (!rc)
        rc = ust_k_assign_materials_field(&ctx->grid, ctx->cell_off, ctx->np, ctx->xi0, ctx->xi1, ctx->xi2,
                                          (const float *)d_f, lo, hi, nbins, first_material, ctx->mat, ctx->stream);
    if (!rc) rc = ust_stream_sync(ctx->stream);
    ust_dev_free(d_f);
    if (rc) UST_FAIL(ctx, rc, "material field assignment failed: %s", ust_dev_error_string());
    ctx->assembled = 0;
    return UST_OK;
}

int ust_swarm_cohesion_noise(ust_context *ctx, float amplitude, unsigned long long seed) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    if (!ctx->coh) DALLOC(ctx, ctx->coh, (size_t)ctx->np * sizeof(float));
    CKD(ctx, ust_k_noise_scale(ctx->coh, ctx->np, amplitude, seed, ctx->stream));
    ctx->assembled = 0;
    return UST_OK;
}

int ust_swarm_download(ust_context *ctx, int *cell_off, double *xi, double *w, unsigned char *material) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    size_t n = (size_t)ctx->np;
    if (cell_off) CKD(ctx, ust_memcpy_d2h(cell_off, ctx->cell_off, (size_t)(ctx->nel + 1) * sizeof(int), ctx->stream));
    if (xi) {
        double *tmp = (double *)malloc(n * sizeof(double));
        if (!tmp) UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
        const double *src[3] = {ctx->xi0, ctx->xi1, ctx->xi2};
        for (int d = 0; d < 3; d++) {
            if (d < ctx->grid.dim) {
                int rc = ust_memcpy_d2h(tmp, src[d], n * sizeof(double), ctx->stream);
                if (rc) {
                    free(tmp);
                    UST_FAIL(ctx, rc, "download failed: %s", ust_dev_error_string());
                }
            } else {
                memset(tmp, 0, n * sizeof(double));
            }
            for (size_t i = 0; i < n; i++) xi[3 * i + d] = tmp[i];
        }
        free(tmp);
    }
    if (w) CKD(ctx, ust_memcpy_d2h(w, ctx->w, n * sizeof(double), ctx->stream));
    if (material) {
        if (ctx->mat) CKD(ctx, ust_memcpy_d2h(material, ctx->mat, n, ctx->stream));
        else memset(material, 0, n);
    }
    return UST_OK;
}

int ust_swarm_advect(ust_context *ctx, double dt, int order, int allow_fallback, int *n_removed, int *n_empty) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    if (order != 1 && order != 2) UST_FAIL(ctx, UST_EINVAL, "order must be 1 or 2 (got %d)", order);
    if (!isfinite(dt)) UST_FAIL(ctx, UST_EINVAL, "dt is not finite");
    const int np = ctx->np, nel = ctx->nel, dim = ctx->grid.dim;
    const size_t n = (size_t)np;
    size_t wb = ust_k_sort_by_cell_work_bytes(np, nel);
    void *key = NULL, *key_s = NULL, *perm = NULL, *perm_t = NULL, *nxi = NULL, *flag = NULL, *work = NULL;
    void *off = NULL, *oxi = NULL, *ow = NULL, *omat = NULL, *ocoh = NULL;
    int rc = UST_OK, hflag = 0;
    if (!rc) rc = ust_dev_malloc(&key, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&key_s, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&perm, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&perm_t, n * sizeof(unsigned) + 8);
    if (!rc) rc = ust_dev_malloc(&nxi, 3 * n * sizeof(double) + 8);
    if (!rc) rc = ust_dev_malloc(&flag, sizeof(double));
    if (!rc) rc = ust_dev_malloc(&work, wb + 8);
    if (!rc) rc = ust_dev_malloc(&off, (size_t)(nel + 1) * sizeof(int));
    double *n0 = (double *)nxi, *n1 = n0 + n, *n2 = n0 + 2 * n;
    if (!rc) rc = ust_k_fill_f64((double *)flag, 0.0, 1, ctx->stream); /* all-zero bits: the int flag reads 0 */
    if (!rc)
        rc = ust_k_advect(&ctx->grid, ctx->cell_off, np, ctx->xi0, ctx->xi1, ctx->xi2, ctx->u, dt, order,
                          allow_fallback, (unsigned *)key, n0, n1, n2, (int *)flag, ctx->stream);
    if (!rc) rc = ust_memcpy_d2h(&hflag, flag, sizeof(int), ctx->stream);
    if (!rc && hflag) {
        snprintf(ctx->err, sizeof(ctx->err),
                 "a corrector step found no velocity (point outside the box) and allow_fallback is off "
                 "(TimeIntegrand.c:324-327)");
        rc = UST_ENOTCONV;
    }
    if (!rc)
        rc = ust_k_sort_by_cell((const unsigned *)key, (unsigned *)key_s, (unsigned *)perm, (unsigned *)perm_t, np,
                                nel, work, wb, ctx->stream);
    if (!rc) rc = ust_k_cell_offsets((const unsigned *)key_s, np, nel, (int *)off, ctx->stream);
    int *off_h = (int *)malloc((size_t)(nel + 1) * sizeof(int));
    if (!off_h && !rc) rc = UST_ENOMEM;
    if (!rc) rc = ust_memcpy_d2h(off_h, off, (size_t)(nel + 1) * sizeof(int), ctx->stream);
    int kept = 0, empty = 0, max_ppc = 0;
    if (!rc) {
        kept = off_h[nel];
        for (int e = 0; e < nel; e++) {
            const int c = off_h[e + 1] - off_h[e];
            if (c == 0) empty++;
            if (c > max_ppc) max_ppc = c;
        }
    }
    free(off_h);
    const size_t k = (size_t)kept;
    if (!rc) rc = ust_dev_malloc(&oxi, 3 * k * sizeof(double) + 8);
    if (!rc) rc = ust_dev_malloc(&ow, k * sizeof(double) + 8);
    if (!rc && ctx->mat) rc = ust_dev_malloc(&omat, k + 8);
    if (!rc && ctx->coh) rc = ust_dev_malloc(&ocoh, k * sizeof(float) + 8);
    double *o0 = (double *)oxi, *o1 = o0 + k, *o2 = o0 + 2 * k;
    if (!rc)
        rc = ust_k_swarm_gather(dim, kept, (const unsigned *)perm, (const unsigned *)key_s, (const int *)off, n0, n1,
                                n2, ctx->mat, ctx->coh, o0, o1, o2, (double *)ow, (unsigned char *)omat,
                                (float *)ocoh, ctx->stream);
    /* the strain-weakening state travels with its particles */
    void *opfw = NULL, *oinc = NULL;
    if (!rc && ctx->pfw) {
        rc = ust_dev_malloc(&opfw, k * sizeof(double) + 8);
        if (!rc) rc = ust_dev_malloc(&oinc, k * sizeof(double) + 8);
        if (!rc) rc = ust_k_permute_f64((const unsigned *)perm, ctx->pfw, (double *)opfw, kept, ctx->stream);
        if (!rc) rc = ust_k_permute_f64((const unsigned *)perm, ctx->pfw_inc, (double *)oinc, kept, ctx->stream);
    }
    if (!rc) {
        rc = ust_swarm_upload_dev(ctx, kept, (const int *)off, o0, o1, dim == 3 ? o2 : NULL, (const double *)ow,
                                  (const unsigned char *)omat, (const float *)ocoh, max_ppc);
        if (!rc && opfw) { /* upload_dev released the old state: the permuted arrays become the state */
            ctx->pfw = (double *)opfw;
            ctx->pfw_inc = (double *)oinc;
            opfw = oinc = NULL;
        }
    } else if (!ctx->err[0] || rc != UST_ENOTCONV) {
        snprintf(ctx->err, sizeof(ctx->err), "swarm advection failed (%d): %s", rc, ust_dev_error_string());
    }
    ust_dev_free(opfw);
    ust_dev_free(oinc);
    void *tmp[] = {key, key_s, perm, perm_t, nxi, flag, work, off, oxi, ow, omat, ocoh};
    for (size_t i = 0; i < sizeof(tmp) / sizeof(tmp[0]); i++) ust_dev_free(tmp[i]);
    if (rc) return rc;
    if (n_removed) *n_removed = np - kept;
    if (n_empty) *n_empty = empty;
    return UST_OK;
}

void ust_pcdvc_default(ust_pcdvc *o) {
    memset(o, 0, sizeof(*o));
    o->res[0] = o->res[1] = o->res[2] = 10;
    o->upper_t = 25.0;
    o->lower_t = 0.6;
    o->max_deletions = 2;
    o->max_splits = 3;
}

int ust_swarm_pcdvc(ust_context *ctx, const ust_pcdvc *opts, int *n_split, int *n_deleted) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    if (!opts) UST_FAIL(ctx, UST_EINVAL, "PCDVC options missing");
    if (opts->inflow || opts->allow_splitting_in_boundary_interface_cells)
        UST_FAIL(ctx, UST_EINVAL, "PCDVC: the Inflow branch (random re-seeding of elements, PCDVC.c:1216-1312) is "
                                  "not supported");
    if (opts->max_splits < 0 || opts->max_deletions < 0 || opts->max_splits > 4096)
        UST_FAIL(ctx, UST_EINVAL, "PCDVC: maxSplits %d / maxDeletions %d out of range", opts->max_splits,
                 opts->max_deletions);
    ust_pcdvc_opts ko;
    ko.upper_t = opts->upper_t;
    ko.lower_t = opts->lower_t;
    if (ko.upper_t < ko.lower_t) { /* the reference resets both (PCDVC.c:270-274) */
        ko.lower_t = 0.6;
        ko.upper_t = 25.0;
    }
    ko.max_deletions = opts->max_deletions;
    ko.max_splits = opts->max_splits;
    ko.split_in_interface_cells = opts->split_in_interface_cells;
    ko.delete_in_interface_cells = opts->delete_in_interface_cells;
    const int dim = ctx->grid.dim, nel = ctx->nel, np = ctx->np;
    /* 1. the regions of the current particles (weights and centroids in the integration points) */
    int rc = ust_swarm_dvc_weights(ctx, opts->res, 1);
    if (rc) return rc;
    /* 2. the per-element plans */
    void *cnt = NULL, *list = NULL, *flag = NULL, *noff = NULL, *o = NULL, *ow = NULL, *omat = NULL, *ocoh = NULL;
    void *os0 = NULL, *os1 = NULL;
    int hflag[4] = {0, 0, 0, 0};
    int *cnt_h = NULL, *off_h = NULL;
    const size_t lcap = (size_t)np + (size_t)nel * (size_t)ko.max_splits;
    if (!rc) rc = ust_dev_malloc(&cnt, (size_t)nel * sizeof(int) + 8);
    if (!rc) rc = ust_dev_malloc(&list, lcap * sizeof(int) + 8);
    if (!rc) rc = ust_dev_malloc(&flag, 2 * sizeof(double));
    if (!rc) rc = ust_k_fill_f64((double *)flag, 0.0, 2, ctx->stream); /* all-zero bits: the ints read 0 */
    if (!rc) rc = ust_k_pcdvc_plan(&ctx->grid, ctx->cell_off, ctx->w, ctx->mat, &ko, (int *)cnt, (int *)list,
                                   (int *)flag, ctx->stream);
    if (!rc) rc = ust_memcpy_d2h(hflag, flag, 3 * sizeof(int), ctx->stream);
    cnt_h = (int *)malloc((size_t)nel * sizeof(int));
    off_h = (int *)malloc((size_t)(nel + 1) * sizeof(int));
    if (!rc && (!cnt_h || !off_h)) rc = UST_ENOMEM;
    if (!rc) rc = ust_memcpy_d2h(cnt_h, cnt, (size_t)nel * sizeof(int), ctx->stream);
    long long total = 0;
    int max_ppc = 0;
    if (!rc) {
        off_h[0] = 0;
        for (int e = 0; e < nel; e++) {
            total += cnt_h[e];
            if (cnt_h[e] > max_ppc) max_ppc = cnt_h[e];
            off_h[e + 1] = total > 2147483647LL ? 0 : (int)total;
        }
        if (total > 2147483647LL) rc = UST_EINVAL;
    }
    if (!rc && (hflag[0] & 1)) {
        free(cnt_h);
        free(off_h);
        ust_dev_free(cnt);
        ust_dev_free(list);
        ust_dev_free(flag);
        UST_FAIL(ctx, UST_EINVAL, "under resolved cell after population control: a changed element keeps fewer than 3 "
                                  "particles (PCDVC.c:1418-1421)");
    }
    /* 3. the new swarm */
    const size_t k = (size_t)total;
    if (!rc) rc = ust_dev_malloc(&noff, (size_t)(nel + 1) * sizeof(int));
    if (!rc) rc = ust_memcpy_h2d(noff, off_h, (size_t)(nel + 1) * sizeof(int), ctx->stream);
    if (!rc) rc = ust_dev_malloc(&o, 3 * k * sizeof(double) + 8);
    if (!rc) rc = ust_dev_malloc(&ow, k * sizeof(double) + 8);
    if (!rc && ctx->mat) rc = ust_dev_malloc(&omat, k + 8);
    if (!rc && ctx->coh) rc = ust_dev_malloc(&ocoh, k * sizeof(float) + 8);
    if (!rc && ctx->pfw) rc = ust_dev_malloc(&os0, k * sizeof(double) + 8);
    if (!rc && ctx->pfw) rc = ust_dev_malloc(&os1, k * sizeof(double) + 8);
    double *o0 = (double *)o, *o1 = o0 + k, *o2 = o0 + 2 * k;
    if (!rc)
        rc = ust_k_pcdvc_build(dim, nel, ko.max_splits, ctx->cell_off, (const int *)noff, (const int *)list, ctx->xi0,
                               ctx->xi1, ctx->xi2, ctx->ixi0, ctx->ixi1, ctx->ixi2, ctx->mat, ctx->coh, ctx->pfw,
                               ctx->pfw_inc, o0, o1, dim == 3 ? o2 : NULL, (unsigned char *)omat, (float *)ocoh,
                               (double *)os0, (double *)os1, ctx->stream);
    if (!rc) rc = ust_k_fill_f64((double *)ow, 0.0, k, ctx->stream);
    if (!rc) {
        rc = ust_swarm_upload_dev(ctx, (int)k, (const int *)noff, o0, o1, dim == 3 ? o2 : NULL, (const double *)ow,
                                  (const unsigned char *)omat, (const float *)ocoh, max_ppc);
        if (!rc && os0) { /* upload_dev released the old state: the rebuilt arrays become the state */
            ctx->pfw = (double *)os0;
            ctx->pfw_inc = (double *)os1;
            os0 = os1 = NULL;
        }
    } else if (!ctx->err[0]) {
        snprintf(ctx->err, sizeof(ctx->err), "PCDVC population control failed (%d): %s", rc, ust_dev_error_string());
    }
    free(cnt_h);
    free(off_h);
    void *tmp[] = {cnt, list, flag, noff, o, ow, omat, ocoh, os0, os1};
    for (size_t i = 0; i < sizeof(tmp) / sizeof(tmp[0]); i++) ust_dev_free(tmp[i]);
    if (rc) return rc;
    /* 4. the regions of the new particles: integration points at the centroids, weights the volumes */
    rc = ust_swarm_dvc_weights(ctx, opts->res, 1);
    if (rc) return rc;
    if (n_split) *n_split = hflag[1];
    if (n_deleted) *n_deleted = hflag[2];
    return UST_OK;
}

int ust_swarm_count(ust_context *ctx) { return ctx->np; }

int ust_swarm_download_integration_points(ust_context *ctx, double *xi, double *w) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    size_t n = (size_t)ctx->np;
    if (xi) {
        double *tmp = (double *)malloc(n * sizeof(double));
        if (!tmp) UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
        const int sep = ctx->ixi0 != NULL;
        const double *src[3] = {sep ? ctx->ixi0 : ctx->xi0, sep ? ctx->ixi1 : ctx->xi1, sep ? ctx->ixi2 : ctx->xi2};
        for (int d = 0; d < 3; d++) {
            if (d < ctx->grid.dim) {
                int rc = ust_memcpy_d2h(tmp, src[d], n * sizeof(double), ctx->stream);
                if (rc) {
                    free(tmp);
                    UST_FAIL(ctx, rc, "download failed: %s", ust_dev_error_string());
                }
            } else {
                memset(tmp, 0, n * sizeof(double));
            }
            for (size_t i = 0; i < n; i++) xi[3 * i + d] = tmp[i];
        }
        free(tmp);
    }
    if (w) CKD(ctx, ust_memcpy_d2h(w, ctx->w, n * sizeof(double), ctx->stream));
    return UST_OK;
}

int ust_swarm_dvc_weights(ust_context *ctx, const int *res, int set_centroids) {
    if (!ctx->np) UST_FAIL(ctx, UST_ESTATE, "no swarm");
    if (!res) UST_FAIL(ctx, UST_EINVAL, "resolution missing");
    const int dim = ctx->grid.dim;
    int ncell = 1;
    for (int a = 0; a < dim; a++) {
        if (res[a] < 1 || res[a] > 64) UST_FAIL(ctx, UST_EINVAL, "DVC resolution %d out of 1..64", res[a]);
        ncell *= res[a];
    }
    if (ncell > 4096) UST_FAIL(ctx, UST_EINVAL, "DVC sub-grid of %d cells exceeds 4096", ncell);
    if (ctx->max_ppc < 1 || ctx->max_ppc > 32767)
        UST_FAIL(ctx, UST_EINVAL, "%d particles per element: DVC holds 1..32767", ctx->max_ppc);
    const size_t lds = ust_k_dvc_lds_bytes(dim, res, ctx->max_ppc);
    if (lds > 65536)
        UST_FAIL(ctx, UST_EINVAL, "DVC needs %zu bytes of LDS per element (%d particles, %d sub-cells): limit 65536",
                 lds, ctx->max_ppc, ncell);
    /* DVCWeights moves the integration points (DVCWeights.c:836-843), not the material points they were
     * mapped from: the centroids go to the context's integration-point coordinates, seeded from the
     * material points as CoincidentMapper maps them, and the material points stay where they are */
    const size_t n = (size_t)ctx->np;
    double *d0 = ctx->xi0, *d1 = ctx->xi1, *d2 = ctx->xi2;
    if (set_centroids) {
        if (!ctx->ixi0) {
            DALLOC(ctx, ctx->ixi0, n * sizeof(double));
            DALLOC(ctx, ctx->ixi1, n * sizeof(double));
            if (dim == 3) DALLOC(ctx, ctx->ixi2, n * sizeof(double));
        }
        CKD(ctx, ust_memcpy_d2d(ctx->ixi0, ctx->xi0, n * sizeof(double), ctx->stream));
        CKD(ctx, ust_memcpy_d2d(ctx->ixi1, ctx->xi1, n * sizeof(double), ctx->stream));
        if (dim == 3) CKD(ctx, ust_memcpy_d2d(ctx->ixi2, ctx->xi2, n * sizeof(double), ctx->stream));
        d0 = ctx->ixi0;
        d1 = ctx->ixi1;
        d2 = ctx->ixi2;
    } else {
        free_integration_points(ctx); /* weights only: the integration points coincide with the material points */
    }
    void *flag = NULL;
    int hflag = 0;
    CKD(ctx, ust_dev_malloc(&flag, sizeof(double)));
    int rc = ust_k_fill_f64((double *)flag, 0.0, 1, ctx->stream); /* all-zero bits: the int flag reads 0 */
    if (!rc)
        rc = ust_k_dvc_weights(dim, res, ctx->nel, ctx->cell_off, ctx->max_ppc, d0, d1, d2, ctx->w, set_centroids,
                               (int *)flag, ctx->stream);
    if (!rc) rc = ust_memcpy_d2h(&hflag, flag, sizeof(int), ctx->stream);
    ust_dev_free(flag);
    if (rc) UST_FAIL(ctx, rc, "DVC weights failed (%d): %s", rc, ust_dev_error_string());
    /* the weights and points changed even when an element failed: nothing assembled stays valid */
    ctx->assembled = 0;
    ctx->cache_valid = 0;
    if (hflag & 1)
        UST_FAIL(ctx, UST_EINVAL, "under resolved cell: an element has no particles, add more particles "
                                  "(DVCWeights.c:787)");
    if (hflag & 2) UST_FAIL(ctx, UST_EINVAL, "an element holds more particles than the swarm's maximum per cell");
    if (hflag & 4) UST_FAIL(ctx, UST_ENOTCONV, "DVC region growth did not settle within its round limit");
    return UST_OK;
}
