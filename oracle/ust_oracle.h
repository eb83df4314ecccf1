/*
 * ust_oracle.h -- FP64 CPU restatement of Underworld 1.x's PIC Stokes hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product (libust.so) never links, loads or calls it.
 *
 * Each function follows one reference routine; the citation is on the
 * definition in ust_oracle.c.  Layout conventions (shared with the product):
 *   - Cartesian mesh, nel[d] elements per axis, node id = i + (nx+1)*(j + (ny+1)*k),
 *     element id = i + nx*(j + ny*k)  (CartesianGenerator.c:976-1050 order).
 *   - element-local node a = ix + 2*iy + 4*iz  (BilinearElementType.c:212-243,
 *     TrilinearElementType.c:249-304).
 *   - velocity dof of node n, component c: n*dim + c (full numbering, BCs kept);
 *     "equation" numbering (BC dofs removed) is built by orc_eqnum.
 *   - particles are cell-sorted: cell_off[e]..cell_off[e+1]-1 belong to element e.
 */
#ifndef UST_ORACLE_H
#define UST_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim;
    int nel[3];
    double xmin[3], xmax[3];
} ORC_Mesh;

/* Rheology kinds (Underworld/Rheology component Type names). */
enum {
    ORC_RH_NONE = 0,
    ORC_RH_MATERIALVISCOSITY = 1, /* p0 = eta0 */
    ORC_RH_ARRHENIUS = 2,         /* p0 = eta0, p1 = activationEnergy, p2 = activationVolume, p3 = referenceTemperature */
    ORC_RH_VISCOSITYLIMITER = 3,  /* p0 = maxViscosity, p1 = minViscosity */
    ORC_RH_VONMISES = 4,          /* p0 = cohesion, p1 = minimumViscosity, p2 = strainRateSoftening(0/1),
                                     p3 = cohesionAfterSoftening, p4 = StrainWeakening(0/1) */
    ORC_RH_DRUCKERPRAGER = 5      /* p0 = cohesion, p1 = frictionCoefficient, p2 = minimumYieldStress,
                                     p3 = minimumViscosity, p4 = strainRateSoftening(0/1),
                                     p5 = cohesionAfterSoftening, p6 = frictionCoefficientAfterSoftening,
                                     p7 = StrainWeakening(0/1) */
};

#define ORC_MAX_RHEOLOGIES 4
typedef struct {
    int type;
    double p[8];
} ORC_Rheology;

typedef struct {
    int nrh;
    ORC_Rheology rh[ORC_MAX_RHEOLOGIES];
    /* BuoyancyForceTerm material extension (BuoyancyForceTerm.c:118-122): "density", "alpha",
     * "referenceTemperature", "beta", "referenceDepth" */
    double density, alpha, ref_temperature, beta, ref_depth;
} ORC_Material;

/* BuoyancyForceTerm component keys (BuoyancyForceTerm.c:139-160) and its optional HydrostaticTerm
 * (HydrostaticTerm.c AssignFromXML keys; the background density of HydrostaticTerm_Density, :290-330) */
typedef struct {
    int vertical_axis;        /* "VerticalAxis" (default 1) */
    int use_direction;        /* "gravityDirection" given */
    double direction[3];
    int adjust;               /* "adjust" */
    int hydrostatic;          /* "HydrostaticTerm" given */
    double hs_upper_density, hs_upper_alpha, hs_lower_density, hs_lower_alpha, hs_height, hs_material_boundary;
    double hs_T0, hs_linear_coefficient, hs_exponential_coefficient1, hs_exponential_coefficient2;
    double hs_v, hs_width, hs_time;  /* the moving-boundary terms and the context's current time */
} ORC_Buoyancy;

/* Per-assembly field context for the rheologies (PpcManager lookups). */
typedef struct {
    const double *temperature;   /* nodal Q1 field or NULL */
    const double *velocity;      /* nodal, n_nodes*dim, or NULL */
    const double *pressure;      /* per element (P0) or NULL */
    const float  *coh_scale;     /* per particle cohesion factor (1+noise) or NULL */
    int previous_solution_exists;
    /* StrainWeakening (StrainWeakening.c): postFailureWeakening per particle (NULL: none; negative values are
     * clamped to 0 in place when read, :491-492), the increment written per particle, and the component's
     * softeningStrain / healingRate; DruckerPrager tensileFailure per particle (NULL: not kept) */
    double *pfw;
    double *pfw_inc;
    unsigned char *tensile;
    double softening_strain, healing_rate;
} ORC_Fields;

/* ---- random particle layout ---- */
void   orc_rng_seed(long seed);
double orc_rng_random(void);
double orc_rng_random_minmax(double mn, double mx);
long   orc_rng_state(void);
/* ppc particles per cell, cells in element order, 3 draws per particle. coords: np*3 */
void orc_layout_random(const ORC_Mesh *m, int ppc, long seed, double *coords);
void orc_layout_random_counts(const ORC_Mesh *m, const int *cell_off, long seed, double *coords);

/* ---- mesh / element type ---- */
int  orc_nnodes(const ORC_Mesh *m);
int  orc_nelements(const ORC_Mesh *m);
void orc_vertex(const ORC_Mesh *m, int node, double *x);
void orc_element_nodes(const ORC_Mesh *m, int e, int *nodes);
void orc_shape_functions(int dim, const double *xi, double *N);
void orc_shape_local_derivs(int dim, const double *xi, double GNi[3][8]);
void orc_global_to_local(const ORC_Mesh *m, int e, const double *x, double *xi);
void orc_shape_global_derivs(const ORC_Mesh *m, int e, const double *xi, double *detJ, double GNx[3][8]);
int  orc_find_element(const ORC_Mesh *m, const double *x);

/* coords (np*3, cell-sorted by the layout) -> xi (np*3) and constant weights (np). */
void orc_map_particles(const ORC_Mesh *m, int np, const double *coords, const int *cell_off,
                       double *xi, double *w);

/* ---- rheology chain (one particle) ---- */
/* p: the particle's index in the per-particle arrays of f */
double orc_particle_viscosity(const ORC_Mesh *m, const ORC_Material *mat, int e, const double *xi,
                              const ORC_Fields *f, float coh_scale, long p, unsigned char *yielded);
/* StrainWeakening time step (StrainWeakening.c:401-426): pfw += dt * inc, then negatives to 0 */
void orc_strain_weakening_step(int np, double *pfw, const double *inc, double dt);
void   orc_eval_viscosity(const ORC_Mesh *m, const ORC_Material *mats, int np, const int *cell_off,
                          const double *xi, const int *mat_of, const ORC_Fields *f,
                          double *eta, unsigned char *yielded);

/* ---- buoyancy (BuoyancyForceTerm_AssembleElement, BuoyancyForceTerm.c:377-425) ---- */
/* density of the integration point at local coordinates xi of element e (T nodal or NULL) */
double orc_particle_density(const ORC_Mesh *m, const ORC_Material *mat, int e, const double *xi, const double *T,
                            const ORC_Buoyancy *b);
void orc_eval_density(const ORC_Mesh *m, const ORC_Material *mats, int np, const int *cell_off, const double *xi,
                      const int *mat_of, const double *T, const ORC_Buoyancy *b, double *rho);
/* f_e of one element from per-particle densities: the force along -VerticalAxis or along gravityDirection,
 * with the "adjust" factor cellLocalVolume / sum of weights (BuoyancyForceTerm.c:362-375, 435-448) */
void orc_element_force(const ORC_Mesh *m, int e, int n, const double *xi, const double *w, const double *rho,
                       double gravity, const ORC_Buoyancy *b, double *fe);

/* ---- element terms ---- */
void orc_element_K(const ORC_Mesh *m, int e, int n, const double *xi, const double *w,
                   const double *eta, double *Ke);
void orc_element_G(const ORC_Mesh *m, int e, int n, const double *xi, const double *w, double *Ge);
void orc_element_f(const ORC_Mesh *m, int e, int n, const double *xi, const double *w,
                   const double *rho, double gravity, double *fe);
double orc_element_uzawa(int dim, const double *Ke, const double *Ge);

/* all elements: Ke (nel*nd*nd), Ge (nel*nd), fe (nel*nd), approxS (nel); any may be NULL */
void orc_assemble_elements(const ORC_Mesh *m, const int *cell_off, const double *xi, const double *w,
                           const double *eta, const double *rho, double gravity,
                           double *Ke, double *Ge, double *fe, double *approxS);

/* ---- global system with BCs removed ---- */
typedef struct {
    int n, m;               /* rows, cols */
    int *rowptr, *col;
    double *val;
} ORC_CSR;

void orc_csr_free(ORC_CSR *A);
void orc_csr_spmv(const ORC_CSR *A, const double *x, double *y);
void orc_csr_spmv_t(const ORC_CSR *A, const double *x, double *y);

/* bc_mask: n_nodes*dim bytes, 1 = Dirichlet; bc_val: n_nodes*dim values. eq: n_nodes*dim -> eq or -1 */
int  orc_eqnum(const ORC_Mesh *m, const unsigned char *bc_mask, int *eq);
/* K (neq x neq), G (neq x nel), f (neq), h (nel) with BC lifting (StiffnessMatrix.c:900-951). */
void orc_assemble_global(const ORC_Mesh *m, const int *eq, int neq, const unsigned char *bc_mask,
                         const double *bc_val, const double *Ke, const double *Ge, const double *fe,
                         ORC_CSR *K, ORC_CSR *G, double *f, double *h);
/* K with no BCs, full dof numbering (fixture comparisons) */
void orc_assemble_K_full(const ORC_Mesh *m, const double *Ke, ORC_CSR *K);

/* ---- BSSCR solve ---- */
typedef struct {
    int    mg_levels;         /* 0 = choose automatically */
    int    smooth_its;        /* SSOR sweeps per pre/post smoothing */
    double inner_rtol;        /* A11 tolerance */
    double outer_rtol;        /* scr tolerance */
    int    max_inner, max_outer;
    int    remove_nullspace;  /* constant pressure */
    int    nthreads;
} ORC_SolveOpts;

typedef struct {
    int outer_its;
    int inner_its_total;
    int inner_solves;
    double stokes_rel_residual;
    double t_setup, t_solve;
} ORC_SolveStats;

/* Solve [K G; G^T 0][u;p] = [f;h] (free dofs). u (neq), p (nel). approxS (nel) is the Schur PC. */
int orc_bsscr_solve(const ORC_Mesh *m, const int *eq, int neq, const unsigned char *bc_mask,
                    const ORC_CSR *K, const ORC_CSR *G, const double *f, const double *h,
                    const double *approxS, const ORC_SolveOpts *opt, double *u, double *p,
                    ORC_SolveStats *st);

/* relative Stokes residual ||[f-Ku-Gp; h-G^T u]|| / ||[f;h]|| */
double orc_stokes_residual(const ORC_CSR *K, const ORC_CSR *G, const double *f, const double *h,
                           const double *u, const double *p);

/* one multigrid-preconditioned CG velocity solve K u = b (used for the CPU baseline timing) */
int orc_velocity_solve(const ORC_Mesh *m, const int *eq, int neq, const unsigned char *bc_mask,
                       const ORC_CSR *K, const double *b, double *u, const ORC_SolveOpts *opt,
                       int *its, double *t_setup, double *t_solve);

void orc_set_threads(int n);

/* ---- DVCWeights (PICellerator/Weights/src/DVCWeights.c:203-925) ---- */
/* One element: np particles pts (np x dim, local coordinates in [-1,1]) on the res sub-grid; writes the region
 * centroids cen (np x dim) and volumes w (np).  Returns 0, or -1 when np < 1 ("under resolved cell"). */
int orc_dvc_cell(int dim, const int *res, int np, const double *pts, double *cen, double *w);
/* A cell-sorted swarm, elements in parallel (OpenMP): xi (np x 3, in/out: centroids when set_centroids) and w.
 * Returns 0, or -(e + 1) for the first element e without particles. */
int orc_dvc_swarm(int dim, const int *res, int nel, const int *cell_off, double *xi, double *w, int set_centroids);

#ifdef __cplusplus
}
#endif
#endif
