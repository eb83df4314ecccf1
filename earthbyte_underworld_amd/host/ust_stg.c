/*
 * ust_stg.c -- the reference's element-term plugin API on top of the batched GPU assembly.
 *
 * StgFEM calls one StiffnessMatrixTerm_AssembleElementFunction per element and term
 * (StiffnessMatrixTerm.h:52-58; caller StiffnessMatrix_AssembleElement, StiffnessMatrix.c:1071-1087).
 * A term object is bound to a ust_context once; the first callback after ust_assemble pulls the
 * whole batch (element matrices from the MFMA contraction, G_e, f_e, approxS) to the host and
 * every callback then adds (+=) its element's block into the caller's elStiffMat, exactly as
 * the reference terms accumulate.  Errors abort with a Journal_Firewall-style message, the
 * reference's error behaviour for these callbacks (StiffnessMatrixTerm.c:260-276).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ust_internal.h"

#define UST_MAX_BOUND_TERMS 64

static struct {
    void *term;
    ust_context *ctx;
    int kind;
} g_bound[UST_MAX_BOUND_TERMS];

int ust_stg_bind_term(void *term, ust_context *ctx, int kind) {
    for (int i = 0; i < UST_MAX_BOUND_TERMS; i++)
        if (g_bound[i].term == term || g_bound[i].term == NULL) {
            g_bound[i].term = term;
            g_bound[i].ctx = ctx;
            g_bound[i].kind = kind;
            return UST_OK;
        }
    return UST_ENOMEM;
}

static ust_context *lookup(void *term, int kind) {
    for (int i = 0; i < UST_MAX_BOUND_TERMS; i++)
        if (g_bound[i].term == term && g_bound[i].kind == kind) return g_bound[i].ctx;
    fprintf(stderr, "Error in %s: term %p (kind %d) is not bound to a ust context\n", __func__, term, kind);
    abort();
}

static void firewall(ust_context *ctx, int rc, const char *fn) {
    if (rc == UST_OK) return;
    fprintf(stderr, "Error in func %s: %s\n", fn, ust_error(ctx));
    abort();
}

static void fill_cache(ust_context *ctx) {
    if (ctx->cache_valid) return;
    int nd = (1 << ctx->grid.dim) * ctx->grid.dim;
    size_t ne = (size_t)ctx->nel;
    free(ctx->ke_cache);
    free(ctx->ge_cache);
    free(ctx->fe_cache);
    free(ctx->as_cache);
    ctx->ke_cache = (double *)malloc(ne * nd * nd * sizeof(double));
    ctx->ge_cache = (double *)malloc(ne * nd * sizeof(double));
    ctx->fe_cache = (double *)malloc(ne * nd * sizeof(double));
    ctx->as_cache = (double *)malloc(ne * sizeof(double));
    if (!ctx->ke_cache || !ctx->ge_cache || !ctx->fe_cache || !ctx->as_cache) {
        fprintf(stderr, "Error in func %s: host allocation of the element batch failed\n", __func__);
        abort();
    }
    firewall(ctx, ust_get_element_matrices(ctx, 0, ctx->nel, ctx->ke_cache), __func__);
    firewall(ctx, ust_get_element_G(ctx, 0, ctx->nel, ctx->ge_cache), __func__);
    firewall(ctx, ust_get_element_f(ctx, 0, ctx->nel, ctx->fe_cache), __func__);
    firewall(ctx, ust_get_approxS(ctx, ctx->as_cache), __func__);
    ctx->cache_valid = 1;
}

static void check_element(ust_context *ctx, int e, const char *fn) {
    if (!ctx->assembled) firewall(ctx, UST_ESTATE, fn);
    if (e < 0 || e >= ctx->nel) {
        fprintf(stderr, "Error in func %s: element %d outside [0,%d)\n", fn, e, ctx->nel);
        abort();
    }
}

/* _ConstitutiveMatrixCartesian_AssembleElement (ConstitutiveMatrixCartesian.c:202-412) */
void ust_ConstitutiveMatrixCartesian_AssembleElement(void *constitutiveMatrix, void *stiffnessMatrix, int e,
                                                     void *sle, void *context, double **elStiffMat) {
    (void)stiffnessMatrix;
    (void)sle;
    (void)context;
    ust_context *ctx = lookup(constitutiveMatrix, UST_TERM_CONSTITUTIVE);
    check_element(ctx, e, __func__);
    fill_cache(ctx);
    int nd = (1 << ctx->grid.dim) * ctx->grid.dim;
    const double *ke = ctx->ke_cache + (size_t)e * nd * nd;
    double *out = elStiffMat[0]; /* ReallocArray2D: contiguous row-major (StiffnessMatrix.c:888) */
    for (int i = 0; i < nd * nd; i++) out[i] += ke[i];
}

/* _GradientStiffnessMatrixTerm_AssembleElement (GradientStiffnessMatrixTerm.c:160-250): nd x 1 */
void ust_GradientStiffnessMatrixTerm_AssembleElement(void *matrixTerm, void *stiffnessMatrix, int e, void *sle,
                                                     void *context, double **elStiffMat) {
    (void)stiffnessMatrix;
    (void)sle;
    (void)context;
    ust_context *ctx = lookup(matrixTerm, UST_TERM_GRADIENT);
    check_element(ctx, e, __func__);
    fill_cache(ctx);
    int nd = (1 << ctx->grid.dim) * ctx->grid.dim;
    const double *ge = ctx->ge_cache + (size_t)e * nd;
    for (int v = 0; v < nd; v++) elStiffMat[v][0] += ge[v];
}

/* _UzawaPreconditionerTerm_AssembleElement (UzawaPreconditionerTerm.c:155-236): 1 x 1 for P0 */
void ust_UzawaPreconditionerTerm_AssembleElement(void *matrixTerm, void *stiffnessMatrix, int e, void *sle,
                                                 void *context, double **elPreconditioner) {
    (void)stiffnessMatrix;
    (void)sle;
    (void)context;
    ust_context *ctx = lookup(matrixTerm, UST_TERM_UZAWA_PC);
    check_element(ctx, e, __func__);
    fill_cache(ctx);
    elPreconditioner[0][0] += ctx->as_cache[e];
}

/* _BuoyancyForceTerm_AssembleElement (BuoyancyForceTerm.c:310-452) */
void ust_BuoyancyForceTerm_AssembleElement(void *forceTerm, void *forceVector, int e, double *elForceVec) {
    (void)forceVector;
    ust_context *ctx = lookup(forceTerm, UST_TERM_BUOYANCY);
    check_element(ctx, e, __func__);
    fill_cache(ctx);
    int nd = (1 << ctx->grid.dim) * ctx->grid.dim;
    const double *fe = ctx->fe_cache + (size_t)e * nd;
    for (int v = 0; v < nd; v++) elForceVec[v] += fe[v];
}
