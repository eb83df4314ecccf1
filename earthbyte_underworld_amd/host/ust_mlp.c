/*
 * ust_mlp.c -- learned Schur-preconditioner scaling (host side).
 *
 * The "uw" Schur preconditioner of BSSCR is the per-element approxS = g^T diag(K_e)^-1 g
 * (UzawaPreconditionerTerm.c:195-212, preconditioner.c:138-166), i.e. an inverse-viscosity
 * pressure mass matrix.  A 27 -> 64 -> 64 -> 1 MLP (9 inputs in 2D) reads the log-viscosity
 * neighbourhood of each pressure cell and predicts log10 of a factor on approxS.  Training
 * targets come from solves of the plain method: the Schur diagonal probed with a 3^dim
 * colouring (one A11 solve per colour), divided by approxS.  Plain BSSCR remains the fallback
 * (ust_solve switches back when the learned run stalls).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ust_internal.h"

#define MLP_HIDDEN 64

static void dfree(void **p) {
    if (*p) ust_dev_free(*p);
    *p = NULL;
}
#define DFREE(x) dfree((void **)&(x))
#define DALLOC(ctx, ptr, bytes) CKD(ctx, ust_dev_alloc_zero(ctx, (void **)&(ptr), (bytes)))

int ust_mlp_param_counts(int dim, int hidden, int *nw, int *nb) {
    int nin = dim == 3 ? 27 : 9;
    if (hidden != MLP_HIDDEN) return UST_EINVAL;
    *nw = hidden * nin + hidden * hidden + hidden;
    *nb = 2 * hidden + 1;
    return UST_OK;
}

void ust_mlp_free(ust_context *ctx) {
    ust_mlp *m = &ctx->mlp;
    DFREE(m->w_master);
    DFREE(m->w_bf16);
    DFREE(m->b);
    DFREE(m->mw);
    DFREE(m->vw);
    DFREE(m->mb);
    DFREE(m->vb);
    DFREE(m->gw);
    DFREE(m->gb);
    memset(m, 0, sizeof(*m));
}

static int mlp_alloc(ust_context *ctx, int hidden) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "the MLP input width follows the mesh dimension: set the mesh first");
    ust_mlp *m = &ctx->mlp;
    int nw, nb;
    if (ust_mlp_param_counts(ctx->grid.dim, hidden, &nw, &nb)) UST_FAIL(ctx, UST_EINVAL, "hidden width must be 64");
    if (m->ready && m->nw == nw) return UST_OK;
    ust_mlp_free(ctx);
    m->hidden = hidden;
    m->nw = nw;
    m->nb = nb;
    m->nin = ctx->grid.dim == 3 ? 27 : 9;
    DALLOC(ctx, m->w_master, nw * sizeof(float));
    DALLOC(ctx, m->w_bf16, nw * sizeof(unsigned short));
    DALLOC(ctx, m->b, nb * sizeof(float));
    DALLOC(ctx, m->mw, nw * sizeof(float));
    DALLOC(ctx, m->vw, nw * sizeof(float));
    DALLOC(ctx, m->mb, nb * sizeof(float));
    DALLOC(ctx, m->vb, nb * sizeof(float));
    DALLOC(ctx, m->gw, nw * sizeof(float));
    DALLOC(ctx, m->gb, nb * sizeof(float));
    m->ready = 1;
    m->step = 0;
    return UST_OK;
}

/* splitmix64 -> uniform(-1,1) for the He-uniform initialisation */
static double rnd(unsigned long long *s) {
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return 2.0 * ((double)(z >> 11) / 9007199254740992.0) - 1.0;
}

int ust_mlp_set_params(ust_context *ctx, int hidden, const float *w, const float *b) {
    int rc = mlp_alloc(ctx, hidden);
    if (rc) return rc;
    ust_mlp *m = &ctx->mlp;
    CKD(ctx, ust_memcpy_h2d(m->w_master, w, m->nw * sizeof(float), ctx->stream));
    CKD(ctx, ust_memcpy_h2d(m->b, b, m->nb * sizeof(float), ctx->stream));
    CKD(ctx, ust_k_mlp_to_bf16(m->w_master, m->w_bf16, m->nw, ctx->stream));
    CKD(ctx, ust_k_fill_f32(m->mw, 0.0f, m->nw, ctx->stream));
    CKD(ctx, ust_k_fill_f32(m->vw, 0.0f, m->nw, ctx->stream));
    CKD(ctx, ust_k_fill_f32(m->mb, 0.0f, m->nb, ctx->stream));
    CKD(ctx, ust_k_fill_f32(m->vb, 0.0f, m->nb, ctx->stream));
    m->step = 0;
    CKD(ctx, ust_stream_sync(ctx->stream));
    return UST_OK;
}

int ust_mlp_get_params(ust_context *ctx, float *w, float *b) {
    ust_mlp *m = &ctx->mlp;
    if (!m->ready) UST_FAIL(ctx, UST_ESTATE, "MLP not initialised");
    if (w) CKD(ctx, ust_memcpy_d2h(w, m->w_master, m->nw * sizeof(float), ctx->stream));
    if (b) CKD(ctx, ust_memcpy_d2h(b, m->b, m->nb * sizeof(float), ctx->stream));
    return UST_OK;
}

int ust_mlp_init(ust_context *ctx, int hidden, unsigned long long seed) {
    int nw, nb;
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "set the mesh first");
    if (ust_mlp_param_counts(ctx->grid.dim, hidden, &nw, &nb)) UST_FAIL(ctx, UST_EINVAL, "hidden width must be 64");
    int nin = ctx->grid.dim == 3 ? 27 : 9;
    float *w = (float *)malloc(nw * sizeof(float)), *b = (float *)calloc(nb, sizeof(float));
    unsigned long long s = seed;
    int off = 0;
    const int fan[3] = {nin, hidden, hidden}, cnt[3] = {hidden * nin, hidden * hidden, hidden};
    for (int l = 0; l < 3; l++) {
        double lim = sqrt(6.0 / fan[l]) * (l == 2 ? 0.1 : 1.0);
        for (int i = 0; i < cnt[l]; i++) w[off + i] = (float)(lim * rnd(&s));
        off += cnt[l];
    }
    int rc = ust_mlp_set_params(ctx, hidden, w, b);
    free(w);
    free(b);
    return rc;
}

static int ensure_feat(ust_context *ctx, int n) {
    int nin = ctx->grid.dim == 3 ? 27 : 9;
    if (ctx->feat_cap >= n) return UST_OK;
    DFREE(ctx->feat);
    DFREE(ctx->mlp_out);
    DFREE(ctx->act1);
    DFREE(ctx->act2);
    int nw, nb;
    ust_mlp_param_counts(ctx->grid.dim, MLP_HIDDEN, &nw, &nb);
    DALLOC(ctx, ctx->feat, (size_t)n * nin * sizeof(float));
    DALLOC(ctx, ctx->mlp_out, (size_t)n * sizeof(float) * 2); /* outputs, then targets */
    DALLOC(ctx, ctx->act1, (size_t)256 * (nw + nb) * sizeof(float));
    DALLOC(ctx, ctx->act2, (size_t)256 * sizeof(float) + 64);
    ctx->feat_cap = n;
    return UST_OK;
}

int ust_mlp_features(ust_context *ctx, float *feat) {
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "assemble first");
    int rc = ensure_feat(ctx, ctx->nel);
    if (rc) return rc;
    int nin = ctx->grid.dim == 3 ? 27 : 9;
    CKD(ctx, ust_k_mlp_features(&ctx->grid, ctx->WK, ctx->feat, ctx->stream));
    if (feat) CKD(ctx, ust_memcpy_d2h(feat, ctx->feat, (size_t)ctx->nel * nin * sizeof(float), ctx->stream));
    return UST_OK;
}

/* learned preconditioner diagonal for the current assembly */
int ust_mlp_schur_pc(ust_context *ctx, double *pc_dev) {
    ust_mlp *m = &ctx->mlp;
    if (!m->ready) UST_FAIL(ctx, UST_ESTATE, "learned Schur PC requested but no MLP parameters are set");
    int rc = ust_mlp_features(ctx, NULL);
    if (rc) return rc;
    CKD(ctx, ust_k_mlp_forward(ctx->feat, ctx->nel, m->nin, m->hidden, m->w_bf16, m->b, ctx->mlp_out, NULL, NULL,
                               ctx->stream));
    CKD(ctx, ust_k_apply_scaling(ctx->approxS, ctx->mlp_out, pc_dev, ctx->nel, ctx->stream));
    return UST_OK;
}

int ust_mlp_predict(ust_context *ctx, int n, const float *feat, float *out) {
    ust_mlp *m = &ctx->mlp;
    if (!m->ready) UST_FAIL(ctx, UST_ESTATE, "MLP not initialised");
    int rc = ensure_feat(ctx, n);
    if (rc) return rc;
    CKD(ctx, ust_memcpy_h2d(ctx->feat, feat, (size_t)n * m->nin * sizeof(float), ctx->stream));
    CKD(ctx, ust_k_mlp_forward(ctx->feat, n, m->nin, m->hidden, m->w_bf16, m->b, ctx->mlp_out, NULL, NULL,
                               ctx->stream));
    CKD(ctx, ust_memcpy_d2h(out, ctx->mlp_out, (size_t)n * sizeof(float), ctx->stream));
    return UST_OK;
}

int ust_mlp_gradients(ust_context *ctx, int n, const float *feat, const float *target, float *grad_w, float *grad_b,
                      float *loss) {
    ust_mlp *m = &ctx->mlp;
    if (!m->ready) UST_FAIL(ctx, UST_ESTATE, "MLP not initialised");
    if (n <= 0) UST_FAIL(ctx, UST_EINVAL, "no samples");
    int rc = ensure_feat(ctx, n);
    if (rc) return rc;
    float *tgt = ctx->mlp_out + n;
    CKD(ctx, ust_memcpy_h2d(ctx->feat, feat, (size_t)n * m->nin * sizeof(float), ctx->stream));
    CKD(ctx, ust_memcpy_h2d(tgt, target, (size_t)n * sizeof(float), ctx->stream));
    CKD(ctx, ust_k_mlp_backward(ctx->feat, tgt, n, m->nin, m->hidden, m->w_bf16, m->b, ctx->act1, ctx->act2, NULL,
                                m->gw, m->gb, ctx->act2 + 256, ctx->stream));
    if (grad_w) CKD(ctx, ust_memcpy_d2h(grad_w, m->gw, m->nw * sizeof(float), ctx->stream));
    if (grad_b) CKD(ctx, ust_memcpy_d2h(grad_b, m->gb, m->nb * sizeof(float), ctx->stream));
    if (loss) CKD(ctx, ust_memcpy_d2h(loss, ctx->act2 + 256, sizeof(float), ctx->stream));
    return UST_OK;
}

/* full-batch samples, mini-batches drawn as contiguous windows of a host-side shuffle */
int ust_mlp_train(ust_context *ctx, int n, const float *feat, const float *target, int epochs, int batch, float lr,
                  unsigned long long seed, float *final_loss) {
    ust_mlp *m = &ctx->mlp;
    if (!m->ready) UST_FAIL(ctx, UST_ESTATE, "MLP not initialised");
    if (n <= 0 || batch <= 0) UST_FAIL(ctx, UST_EINVAL, "bad sample or batch count");
    int nin = m->nin, rc;
    if ((rc = ensure_feat(ctx, n))) return rc;
    /* shuffled copy of the training set on the device */
    float *sf = (float *)malloc((size_t)n * nin * sizeof(float)), *st = (float *)malloc((size_t)n * sizeof(float));
    int *perm = (int *)malloc((size_t)n * sizeof(int));
    if (!sf || !st || !perm) {
        free(sf); free(st); free(perm);
        UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    }
    unsigned long long s = seed ? seed : 1;
    float loss = 0.0f;
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = n - 1; i > 0; i--) {
            int j = (int)((rnd(&s) * 0.5 + 0.5) * (i + 1));
            if (j > i) j = i;
            int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
        }
        for (int i = 0; i < n; i++) {
            memcpy(&sf[(size_t)i * nin], &feat[(size_t)perm[i] * nin], nin * sizeof(float));
            st[i] = target[perm[i]];
        }
        float *tgt = ctx->mlp_out + n;
        rc = ust_memcpy_h2d(ctx->feat, sf, (size_t)n * nin * sizeof(float), ctx->stream);
        if (!rc) rc = ust_memcpy_h2d(tgt, st, (size_t)n * sizeof(float), ctx->stream);
        for (int b0 = 0; !rc && b0 < n; b0 += batch) {
            int nb = n - b0 < batch ? n - b0 : batch;
            rc = ust_k_mlp_backward(ctx->feat + (size_t)b0 * nin, tgt + b0, nb, nin, m->hidden, m->w_bf16, m->b,
                                    ctx->act1, ctx->act2, NULL, m->gw, m->gb, ctx->act2 + 256, ctx->stream);
            m->step++;
            if (!rc)
                rc = ust_k_mlp_adam(m->w_master, m->w_bf16, m->b, m->gw, m->gb, m->mw, m->vw, m->mb, m->vb, m->nw,
                                    m->nb, lr, 0.9f, 0.999f, 1e-8f, m->step, ctx->stream);
        }
        if (rc) break;
    }
    free(sf); free(st); free(perm);
    if (rc) UST_FAIL(ctx, rc, "training failed: %s", ust_dev_error_string());
    /* full-set loss with the final weights */
    float *tgt = ctx->mlp_out + n;
    CKD(ctx, ust_memcpy_h2d(ctx->feat, feat, (size_t)n * nin * sizeof(float), ctx->stream));
    CKD(ctx, ust_memcpy_h2d(tgt, target, (size_t)n * sizeof(float), ctx->stream));
    CKD(ctx, ust_k_mlp_backward(ctx->feat, tgt, n, nin, m->hidden, m->w_bf16, m->b, ctx->act1, ctx->act2, NULL, m->gw,
                                m->gb, ctx->act2 + 256, ctx->stream));
    CKD(ctx, ust_memcpy_d2h(&loss, ctx->act2 + 256, sizeof(float), ctx->stream));
    if (final_loss) *final_loss = loss;
    return UST_OK;
}

/* target[e] = log10(diag(S)_e / approxS_e), diag(S) probed with the 3^dim colouring */
int ust_mlp_probe_target(ust_context *ctx, const ust_solve_opts *opts, int nprobe, unsigned long long seed,
                         float *target) {
    (void)nprobe;
    (void)seed;
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "assemble first");
    ust_solve_opts o;
    if (opts) o = *opts;
    else ust_solve_opts_default(&o);
    int rc = ust_solver_prepare(ctx, &o);
    if (rc) return rc;
    int ncol = ctx->grid.dim == 3 ? 27 : 9;
    size_t ne = (size_t)ctx->nel;
    if ((rc = ensure_feat(ctx, ctx->nel))) return rc;
    for (int c = 0; c < ncol; c++) {
        CKD(ctx, ust_k_color_indicator(&ctx->grid, c, ctx->pd, ctx->stream));
        /* q = G^T K^-1 G v */
        CKD(ctx, ust_k_apply_G(&ctx->grid, ctx->WG, ctx->mask_d, ctx->pd, ctx->vb, ctx->stream));
        ust_solve_stats si;
        memset(&si, 0, sizeof(si));
        rc = ust_a11_solve_dev(ctx, &o, ctx->vb, ctx->uprev, &si);
        if (rc) return rc;
        CKD(ctx, ust_k_apply_Gt(&ctx->grid, ctx->WG, ctx->mask_d, ctx->uprev, ctx->pq, 0, ctx->stream));
        CKD(ctx, ust_k_color_gather(&ctx->grid, c, ctx->pq, ctx->pz, ctx->stream));
    }
    CKD(ctx, ust_k_log_ratio(ctx->pz, ctx->approxS, ctx->mlp_out, ctx->nel, ctx->stream));
    CKD(ctx, ust_memcpy_d2h(target, ctx->mlp_out, ne * sizeof(float), ctx->stream));
    return UST_OK;
}
