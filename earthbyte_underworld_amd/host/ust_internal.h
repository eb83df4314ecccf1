/* Internal state of a ust_context (C host layer). */
#ifndef UST_INTERNAL_H
#define UST_INTERNAL_H
#include <stddef.h>

#include "../../include/ust.h"

#define UST_MAX_LEVELS 12
#define UST_MAX_CHEBY 24

/* device scalar slots */
enum { SC_RZ = 0, SC_RZOLD, SC_PQ, SC_RR, SC_BB, SC_SUM, SC_T0, SC_T1, SC_T2, SC_T3,
       SC_O_RZ, SC_O_RZOLD, SC_O_PQ, SC_O_RR,
       SC_CG = 16, /* 4 slots: device-controlled inner CG (threshold, state, iterations, beta) */
       SC_COUNT = 32 };

typedef struct {
    ust_grid g;
    int nnodes, ndof, nel;
    double *WK64;          /* level 0 borrows ctx->WK */
    float *WK32;           /* FP32 V-cycle weights (vcycle_bf16 = 0) */
    unsigned short *WK16;  /* bfloat16 V-cycle weights (vcycle_bf16 = 1) */
    unsigned char *mask;   /* device */
    unsigned char *mask_h; /* host */
    double *invd64;
    float *invd32;
    double lmax;
    int cheb_deg;
    double c1[UST_MAX_CHEBY], c2[UST_MAX_CHEBY];
    /* work vectors */
    double *x64, *b64, *ax64, *d64;
    float *x32, *b32, *ax32, *d32;
    float *x32b;           /* coarse 3D levels: the second x of the out-of-place FP32 smoothing passes */
} ust_level;

typedef struct {
    int hidden, nw, nb, nin;
    float *w_master;        /* device FP32 master copy */
    unsigned short *w_bf16; /* device BF16 weights used by the kernels */
    float *b;               /* device biases */
    float *mw, *vw, *mb, *vb;
    float *gw, *gb;
    int step;
    int ready;
} ust_mlp;

struct ust_context {
    int device;
    void *stream;
    char err[512];
    /* mesh */
    int has_mesh;
    ust_grid grid;
    int nnodes, ndof, nel;
    /* materials */
    int nmat;
    ust_material mats_h[UST_MAX_MATERIALS];
    ust_material *mats_d;
    /* swarm (device SoA).  xi0..2 are the material points' local coordinates (MaterialPointsSwarm: what
     * advection moves and material shapes are evaluated at); w is the integration weight.  The integration
     * points coincide with the material points (CoincidentMapper) unless a weights calculator moved them:
     * then ixi0..2 hold their own local coordinates (DVCWeights' centroids), NULL otherwise. */
    int np, max_ppc;
    int *cell_off;
    double *xi0, *xi1, *xi2, *w;
    double *ixi0, *ixi1, *ixi2;
    unsigned char *mat;
    float *coh;
    double *eta;
    unsigned char *yielded;
    unsigned char *tensile;  /* DruckerPrager tensileFailure per particle */
    double *pfw, *pfw_inc;   /* StrainWeakening postFailureWeakening and its increment per particle (when on) */
    int sw_on;
    ust_strain_weakening sw;
    /* fields */
    double *T;  /* nodal */
    double *u;  /* nodal solution incl. BC values */
    double *p;  /* element */
    int prev_solution;
    /* BCs */
    int has_bcs;
    int bc_nonzero; /* some Dirichlet value is not zero (the RHS needs lifting) */
    unsigned char *mask_d, *mask_h;
    double *bcval_d, *bcval_h;
    int nullspace; /* constant pressure found in the null space of G (ksp_pressure_nullspace.c:139-158 test) */
    double ns_ratio; /* ||G 1|| / ||1|| / ||D||_inf of the last assembly */
    double *tmpv, *tmpp; /* velocity- and pressure-sized scratch of the assembly */
    void *kwork;         /* K u workspace (per-tile node partials) in FP64 elements, for every level's grid */
    size_t kwork_len;
    /* assembly */
    int assembled;
    double gravity;
    ust_buoyancy buoy;   /* BuoyancyForceTerm options (ust_set_buoyancy; default: VerticalAxis 1, nothing else) */
    double *WK, *WG, *WF;
    double *approxS;
    double *f; /* nodal RHS after lifting, zero on BC dofs */
    double *h; /* continuity RHS after lifting */
    /* multigrid hierarchy */
    int levels_ready, levels_prec32, nlev;
    int levels_struct, levels_deg, levels_cdeg, levels_ccdeg, levels_req, coarse_dense, levels_w16;
    double levels_ratio;
    float *WK32_0; /* FP32 copy of the fine weights for ust_apply_K_f32 without a hierarchy */
    int wk32_0_valid;
    ust_level lev[UST_MAX_LEVELS];
    double *Ainv;
    int nc;
    /* solver work */
    double *partial, *scal, *scal_h;
    double *vr, *vz, *vp, *vq, *vt;   /* FP64 velocity work (inner PCG) */
    double *vx;                       /* inner PCG iterate */
    void *pcg_exec[3];                /* captured inner PCG iteration (FP64 / FP32 V-cycle / FP32 PCG), NULL if stale */
    int use_graphs;                   /* 0: launch the inner iteration kernel by kernel (UST_NO_GRAPH=1) */
    int pcg_hint;                     /* iterations of the last inner solve: first convergence check */
    int pcg32_hint[16];               /* the same per refinement step of the FP32 velocity PCG */
    const double *b64_src;            /* set while a level-0 FP32 V-cycle starts from an FP64 right-hand side */
    float *wr, *wz, *wp, *wq, *wx;    /* FP32 velocity work (mixed inner PCG) */
    double *va, *vb;                  /* FP64 velocity work (Schur / IR) */
    double *pr, *pz, *pd, *pq, *ph, *ppc, *pprev; /* pressure work */
    double *uprev;
    double *upic;                     /* Picard: the previous iterate's nodal velocity */
    /* learned scaling */
    ust_mlp mlp;
    float *feat, *mlp_out, *act1, *act2;
    float *user_scale; /* UST_SCHUR_SCALED: log10 factors on approxS (nel) */
    int feat_cap;
    /* element-matrix cache for the per-element plugin callbacks */
    double *ke_cache, *ge_cache, *fe_cache, *as_cache;
    int cache_valid;
};

#define UST_FAIL(ctx, code, ...)                                      \
    do {                                                              \
        snprintf((ctx)->err, sizeof((ctx)->err), __VA_ARGS__);        \
        return (code);                                                \
    } while (0)

#define CKD(ctx, call)                                                                                 \
    do {                                                                                               \
        int _rc = (call);                                                                              \
        if (_rc != UST_OK) {                                                                           \
            snprintf((ctx)->err, sizeof((ctx)->err), "%s failed (%d): %s [%s:%d]", #call, _rc,         \
                     ust_dev_error_string(), __FILE__, __LINE__);                                     \
            return _rc;                                                                                \
        }                                                                                              \
    } while (0)

int ust_levels_setup(ust_context *ctx, const ust_solve_opts *o, int need32);
void ust_levels_free(ust_context *ctx);
void ust_graphs_reset(ust_context *ctx);
int ust_read_scalar(ust_context *ctx, int slot, double *v);
/* the unit direction of the body force: gHat when given, else -1 on the vertical axis */
void ust_buoyancy_direction(const ust_context *ctx, double dir[3]);
int ust_mlp_schur_pc(ust_context *ctx, double *pc_dev);
void ust_mlp_free(ust_context *ctx);
int ust_dev_alloc_zero(ust_context *ctx, void **p, size_t bytes);
/* multigrid levels and work vectors for solves with options o (after an assembly) */
int ust_solver_prepare(ust_context *ctx, const ust_solve_opts *o);
/* K x = b on device vectors (levels and work vectors must be set up) */
int ust_a11_solve_dev(ust_context *ctx, const ust_solve_opts *o, const double *b, double *x, ust_solve_stats *st);

#endif
