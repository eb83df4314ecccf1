/*
 * ust.h -- C host API of the MI355X-native PIC Stokes hot path (Underworld 1.x drop-in).
 *
 * The drop-in boundary.  Every entry point is extern "C", takes plain pointers and sizes and
 * returns an int status (UST_OK = 0, negative on error; ust_error() gives the text).  One
 * context drives one GPU through one HIP stream; a context is not re-entrant (like the
 * reference's StiffnessMatrixTerm objects, StiffnessMatrixTerm.c:247-258, which keep scratch
 * on self).  Host pointers unless the name ends in _dev (device pointers).
 *
 * Which reference interface each entry replaces is stated on the declaration.
 */
#ifndef UST_H
#define UST_H
#include "ust_kernels.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ust_context ust_context;

#define UST_PREC_FP64 0   /* plain: FP64 operators everywhere */
#define UST_PREC_MIXED 1  /* FP32 multigrid V-cycle inside FP64 inner PCG; outer Krylov, residuals and coarse solve FP64 */

#define UST_SCHUR_UW 0       /* approxS = g^T diag(K)^-1 g per element ("-Q22_pc_type uw") */
#define UST_SCHUR_LEARNED 1  /* approxS scaled by the MLP; plain "uw" remains the fallback */
#define UST_SCHUR_SCALED 2   /* approxS scaled by 10^s, s set by ust_set_schur_scaling; plain "uw" the fallback */
#define UST_SCHUR_GKGDIAG 3  /* diag(G^T diag(K)^-1 G) with the assembled diagonal of K ("-Q22_pc_type gkgdiag",
                                BSSCR_FormSchurApproximationDiag, preconditioner.c:268-310) */

typedef struct {
    int precision;           /* UST_PREC_* */
    int schur_pc;            /* UST_SCHUR_* */
    int mg_levels;           /* 0: as many as the grid allows */
    int smoother_degree;     /* Chebyshev degree of each pre/post smoothing (default 2) */
    double cheby_ratio;      /* lambda_max / lambda_min of the Chebyshev interval (default 12) */
    double outer_rtol;       /* "-scr_ksp_rtol" */
    double inner_rtol;       /* "-A11_ksp_rtol" */
    int max_outer;           /* "-scr_ksp_max_it" */
    int max_inner;           /* "-A11_ksp_max_it" */
    int remove_nullspace;    /* -1: when ||G 1|| / ||1|| < 1e-7 ||D||_inf (ksp_pressure_nullspace.c:139-158); 0/1 force */
    int stall_window;        /* learned PC: fall back to plain if the residual falls by less than 10 % over this
                                many iterations */
    int inner_relax;         /* 1: the inner tolerance of each Schur iteration is inner_rtol * ||h_hat|| / ||r_k||
                                (capped at inner_relax_cap), relaxed as the outer residual falls; 0: fixed, as
                                the reference */
    double inner_relax_cap;  /* largest relaxed inner tolerance (default 1e-2) */
    double stokes_rtol;      /* > 0: after the back-solve, while the FP64 Stokes residual exceeds this, restart the
                                Schur CG from the current pressure on its true residual with a 10x tighter outer
                                tolerance (at most 3 restarts); 0: stop at outer_rtol, as the reference */
    int verbose;
    int coarse_degree;       /* Chebyshev degree on the multigrid levels between the finest and the coarsest
                                (0: smoother_degree) */
    int coarsest_degree;     /* Chebyshev degree on the coarsest level when it is too large for the dense
                                solve (0: 24) */
    int vcycle_bf16;         /* mixed precision: 1 (default) the FP32 V-cycle reads bfloat16 K weights (half the
                                weight stream of every multigrid pass), 0 FP32 weights */
    int inner_fp32;          /* mixed precision: 1 runs the velocity PCG itself in FP32 (FP32 K weights and Krylov
                                vectors, dots accumulated in FP64) inside FP64 iterative refinement: each refinement
                                solves for the correction of the FP64 residual, rounded to FP32 at unit norm, to the
                                remaining tolerance (at most 1e-5 per step), and the FP64 residual is recomputed only
                                when the requested tolerance is tighter than 1e-4; 0 (default) FP64 PCG around the
                                FP32 V-cycle */
} ust_solve_opts;

typedef struct {
    int outer_its;
    int inner_its;           /* total inner (velocity) Krylov iterations */
    int inner_solves;
    int refinements;         /* FP64 iterative-refinement steps (mixed) */
    int fallback_used;       /* learned PC stalled and plain BSSCR finished the solve */
    int mg_levels;
    double stokes_residual;  /* ||[f-Ku-Gp; h-G^T u]|| / ||[f;h]|| in FP64 */
    double momentum_residual, continuity_residual;
    double t_assemble_ms, t_setup_ms, t_solve_ms;
    int restarts;            /* Schur CG restarts taken for stokes_rtol */
    int nullspace_removed;   /* the constant pressure was projected out (ksp_pressure_nullspace.c test) */
    double nullspace_ratio;  /* ||G 1|| / ||1|| / ||D||_inf of the assembly (removed when < 1e-7) */
    int converged;           /* 1: the stopping test held (stokes_rtol when set, else outer_rtol); 0: max_outer
                                ended the Schur CG first (the solve still returns UST_OK, as KSP at its limit) */
    int hhat_rescaled;       /* ||h_hat|| / ||G||_inf was in (1e-20, 1e-6): the Schur solve ran on h_hat scaled by
                                its inverse and the pressure was scaled back (auglag-driver-DGTGD.c:333-358) */
} ust_solve_stats;

void ust_solve_opts_default(ust_solve_opts *o);

/* ---- lifecycle ---- */
int ust_create(ust_context **ctx, int device);
void ust_destroy(ust_context *ctx);
const char *ust_error(const ust_context *ctx);
void *ust_get_stream(ust_context *ctx);

/* ---- mesh: CartesianGenerator "elementResI/J/K", "minX..maxZ" (CartesianGenerator.c:976-1050) ---- */
int ust_set_mesh(ust_context *ctx, int dim, const int nel[3], const double xmin[3], const double xmax[3]);

/* ---- materials: RheologyMaterial "Rheology" list, in registration order (RheologyMaterial.c:240-270) ---- */
int ust_set_materials(ust_context *ctx, int nmat, const ust_material *mats);
/* XML component Type + keys -> rheology (Arrhenius.c:110-133, MaterialViscosity.c:115,
 * ViscosityLimiter.c:101-103, VonMises.c:157-159, YieldRheology.c:140-145, DruckerPrager.c:173-189).  The
 * "StrainWeakening" key takes 1 when the rheology names the context's StrainWeakening component. */
int ust_rheology_from_xml(const char *type_name, int nkeys, const char *const *keys, const double *values,
                          ust_rheology *out);

/* ---- swarm: IntegrationPointsSwarm cell tables (cellParticleTbl/cellParticleCountTbl) ----
 * cell_off (nel+1), xi (np*3, local coords), w (np), material (np, may be NULL = 0),
 * coh_scale (np, may be NULL = 1): per-particle cohesion factor. */
int ust_swarm_upload(ust_context *ctx, int np, const int *cell_off, const double *xi, const double *w,
                     const unsigned char *material, const float *coh_scale);
/* Same from device SoA arrays (copied). */
int ust_swarm_upload_dev(ust_context *ctx, int np, const int *cell_off_dev, const double *xi0_dev,
                         const double *xi1_dev, const double *xi2_dev, const double *w_dev,
                         const unsigned char *material_dev, const float *coh_scale_dev, int max_ppc);

/* Swarm generated on the device: RandomParticleLayout + coincident mapping + ConstantWeights.
 * cell_off (host, nel+1) gives the particle count of every cell (NULL: ppc in every cell). */
int ust_swarm_random(ust_context *ctx, const int *cell_off, int ppc, long seed);
/* SpaceFillerParticleLayout, the material swarm layout of the reference's PIC components
 * (PICellerator/Apps/PIC_Components/MaterialPointSwarm.xml; SpaceFillerParticleLayout.c:218-240,
 * GlobalParticleLayout.c:160-318): (unsigned)(nel x avg_ppc) particles at the points of the first dim Sobol
 * sequences of SobolGenerator's table scaled to the box, each in the element Mesh_RegularAlgorithms_SearchElements
 * gives it (faces to the lower element), in generation order within an element; coincident mapping and
 * ConstantWeights.  Elements may be left empty (population control or DVC then decide). */
int ust_swarm_spacefiller(ust_context *ctx, double avg_ppc);
/* material index per particle from cell_material (host, nel; may be NULL) and shapes in order */
int ust_swarm_assign_materials(ust_context *ctx, int nshape, const ust_shape *shapes,
                               const unsigned char *cell_material);
/* material per particle from a nodal field (host, float per node) sampled at the particle and quantised to
 * nbins materials first_material.. over [lo, hi] (ust_k_assign_materials_field) */
int ust_swarm_assign_materials_field(ust_context *ctx, const float *field_nodal, double lo, double hi, int nbins,
                                     int first_material);
/* per-particle cohesion factor 1 + amplitude*u, u uniform in (-1,1) (seeded) */
int ust_swarm_cohesion_noise(ust_context *ctx, float amplitude, unsigned long long seed);
/* copy the swarm back (any pointer may be NULL): cell_off (nel+1), material-point local coordinates xi (np*3),
 * integration weights w (np), material (np) */
int ust_swarm_download(ust_context *ctx, int *cell_off, double *xi, double *w, unsigned char *material);
/* Advect the swarm through the current nodal velocity (the last solve's) by dt and re-sort it by element
 * (SwarmAdvector + TimeIntegrator + ParticleMovementHandler + ConstantWeights): order 1 (Euler) or 2
 * (midpoint, the TimeIntegrator default); allow_fallback: a corrector point outside the box takes the first-
 * order step instead of failing (allowFallbackToFirstOrder).  Particles that end outside the box are removed;
 * *n_removed counts them, *n_empty the elements left without particles (their element terms are then zero:
 * population control is not part of this step). */
int ust_swarm_advect(ust_context *ctx, double dt, int order, int allow_fallback, int *n_removed, int *n_empty);
/* DVCWeights as the swarm's weights calculator (PICellerator/Weights/src/DVCWeights.c:763-925, replacing
 * WeightsCalculator_CalculateAll's per-cell _DVCWeights_Calculate): the particles of each element grow discrete
 * Voronoi regions on a res[0] x res[1] (x res[2]) sub-grid of the local element (XML resolutionX/Y/Z, default
 * 10), each weight becomes its region's volume and, when set_centroids (the reference's behaviour), the
 * integration points move to the region centroids (DVCWeights.c:836-843).  The material points stay where they
 * are: the context keeps the integration points' own coordinates, which the assembly uses, until the swarm
 * changes (upload, layout or advection map the integration points onto the material points again, with
 * ConstantWeights; call this again after each advection, as the reference recalculates the weights after each
 * remap).  An element without particles is an error ("under resolved cell", DVCWeights.c:787). */
int ust_swarm_dvc_weights(ust_context *ctx, const int *res, int set_centroids);
int ust_swarm_count(ust_context *ctx);
/* the integration points (IntegrationPointsSwarm: local coordinates xi (np*3) and weights w (np), either may
 * be NULL): the material points' coordinates unless ust_swarm_dvc_weights moved them to the centroids */
int ust_swarm_download_integration_points(ust_context *ctx, double *xi, double *w);
/* per-particle cohesion factors (np floats; 1 where none were set) */
int ust_get_cohesion_scale(ust_context *ctx, float *coh);

/* ---- BuoyancyForceTerm options (BuoyancyForceTerm.c:139-160; material keys density, alpha,
 * referenceTemperature, beta, referenceDepth are ust_material fields): VerticalAxis, gravityDirection,
 * adjust, HydrostaticTerm.  NULL restores the defaults (VerticalAxis 1, force along -y). ---- */
int ust_set_buoyancy(ust_context *ctx, const ust_buoyancy *b);

/* ---- PCDVC population control (PICellerator/Utils/src/PCDVC.c) ---- */
typedef struct {
    int res[3];                 /* DVCWeights "resolutionX/Y/Z" (default 10) */
    double upper_t, lower_t;    /* "upperT" (25), "lowerT" (0.6), percent of the element volume */
    int max_deletions;          /* "maxDeletions" (2) */
    int max_splits;             /* "maxSplits" (3) */
    int split_in_interface_cells;  /* "splitInInterfaceCells" (default Inflow, 0) */
    int delete_in_interface_cells; /* "deleteInInterfaceCells" (0) */
    int allow_splitting_in_boundary_interface_cells; /* (0); 1 turns boundary elements into Inflow elements */
    int inflow;                 /* "Inflow" (0); the Inflow branch re-seeds elements with rand() and is not built */
} ust_pcdvc;
void ust_pcdvc_default(ust_pcdvc *o);
/* _PCDVC_Calculate over every element (PCDVC.c:1493-1523): DVC weights, then splits of particles whose region
 * exceeds upperT % of the element (a copy at the region centroid, with the parent's material and state) and
 * deletions of particles whose region is below lowerT %, then DVC again; the integration points end at the
 * region centroids with the region volumes as weights.  Deletion candidates are taken in decreasing swarm
 * position within the element.  Returns the numbers of splits and deletions; UST_EINVAL for the Inflow
 * options, for an element without particles, and for a changed element left with fewer than 3 particles. */
int ust_swarm_pcdvc(ust_context *ctx, const ust_pcdvc *opts, int *n_split, int *n_deleted);

/* ---- StrainWeakening component (StrainWeakening.c:217-280 keys, defaults as there) ---- */
typedef struct {
    double softening_strain;          /* "softeningStrain" (HUGE_VAL) */
    double healing_rate;              /* "healingRate" (0) */
    double initial_damage_fraction;   /* "initialDamageFraction" (0) */
    double initial_damage_wavenumber; /* "initialDamageWavenumber" (-1: off) */
    double wavenumber_sin[3];         /* "initialDamageWavenumberSinI/J/K" (-1: off) */
    double wavenumber_cos[3];         /* "initialDamageWavenumberCosI/J/K" (-1: off) */
    double initial_damage_factor;     /* "initialDamageFactor" (1) */
    long random_seed;                 /* "randomSeed" (0) */
    int has_shape;                    /* "initialStrainShape" given */
    ust_shape initial_strain_shape;
} ust_strain_weakening;
void ust_strain_weakening_default(ust_strain_weakening *sw);
/* Turns the per-particle weakening state on (sw) or off (NULL).  Yield rheologies whose StrainWeakening slot
 * is set read it (VonMises.c:194-199, DruckerPrager.c:283-291) and every assembly with a previous solution
 * writes the increment (YieldRheology.c:224-226).  The state starts at 0 for every particle. */
int ust_set_strain_weakening(ust_context *ctx, const ust_strain_weakening *sw);
/* _StrainWeakening_Initialise (StrainWeakening.c:318-399): the initial damage of every particle, in swarm order,
 * from the C library's srand(randomSeed)/rand() sequence as the reference draws it, modulated by the
 * wavenumber factors and restricted to initialStrainShape; increments set to 0. */
int ust_strain_weakening_initialise(ust_context *ctx);
/* one TimeIntegrand step of postFailureWeakening: += dt * increment, then negatives to 0
 * (StrainWeakening.c:401-426) */
int ust_strain_weakening_step(ust_context *ctx, double dt);
/* per-particle state in swarm order (np each; either pointer may be NULL) */
int ust_get_strain_weakening(ust_context *ctx, double *pfw, double *pfw_inc);
int ust_set_strain_weakening_state(ust_context *ctx, const double *pfw, const double *pfw_inc);
/* DruckerPrager tensileFailure flag of every particle from the last assembly with a previous solution
 * (DruckerPrager.c:312) */
int ust_get_tensile_failure(ust_context *ctx, unsigned char *tensile);

/* ---- fields read by the rheologies (PpcManager lookups, PpcManager.c:503-606) ---- */
int ust_set_temperature(ust_context *ctx, const double *T_nodal);
/* velocity (nodal, n_nodes*dim, BC values included) and pressure (per element) used by yield */
int ust_set_solution(ust_context *ctx, const double *u_nodal, const double *p);
int ust_set_previous_solution_exists(ust_context *ctx, int flag);

/* ---- boundary conditions: mask (n_nodes*dim, 1 = Dirichlet) and values ---- */
int ust_set_bcs(ust_context *ctx, const unsigned char *mask, const double *values);

/* ---- assembly (replaces the K, G, preconditioner and force assembly of the Stokes SLE:
 * __StiffnessMatrix_NewAssemble StiffnessMatrix.c:825-1017 with ConstitutiveMatrixCartesian,
 * GradientStiffnessMatrixTerm, UzawaPreconditionerTerm, ForceVector_Assemble + BuoyancyForceTerm) ---- */
int ust_assemble(ust_context *ctx, double gravity);

/* batched element results (drop-in for the per-element elStiffMat callbacks) */
int ust_get_element_matrices(ust_context *ctx, int first, int count, double *Ke); /* count*nd*nd */
int ust_get_element_matrices_dev(ust_context *ctx, int first, int count, double *Ke_dev);
int ust_get_element_G(ust_context *ctx, int first, int count, double *Ge);         /* count*nd */
int ust_get_element_f(ust_context *ctx, int first, int count, double *fe);         /* count*nd */
int ust_get_approxS(ust_context *ctx, double *approxS);                            /* nel */
int ust_get_particle_viscosity(ust_context *ctx, double *eta, unsigned char *yielded);
int ust_get_weights(ust_context *ctx, double *WK, double *WG, double *WF);         /* SoA (nq|nv)*nel */
/* device views for callers that drive the kernel ABI directly (valid until the next ust_set_mesh) */
int ust_get_device_views(ust_context *ctx, ust_grid *grid, const double **WK, const double **WG,
                         const double **WF, const unsigned char **mask);
/* right-hand sides after BC lifting: f (nodal, zero on BC dofs), h (per element) */
int ust_get_rhs(ust_context *ctx, double *f_nodal, double *h);

/* ---- operators (MatMult inside MatSchurComplement / PCMG), nodal vectors, BC dofs = 0 ---- */
int ust_apply_K(ust_context *ctx, const double *u, double *y);
int ust_apply_K_dev(ust_context *ctx, const double *u_dev, double *y_dev);
int ust_apply_K_f32_dev(ust_context *ctx, const float *u_dev, float *y_dev);
int ust_apply_K_f32(ust_context *ctx, const float *u, float *y);
int ust_apply_G(ust_context *ctx, const double *p, double *y);
int ust_apply_Gt(ust_context *ctx, const double *u, double *q);

/* ---- BSSCR solve (KSPSolve_BSSCR BSSCR.c:151-229 / BSSCR_DRIVER_auglag auglag-driver-DGTGD.c:53-441) ---- */
int ust_solve(ust_context *ctx, const ust_solve_opts *opts, ust_solve_stats *stats);
int ust_get_solution(ust_context *ctx, double *u_nodal, double *p);
/* the diagonal Schur preconditioner of kind schur_pc (UST_SCHUR_*) that ust_solve with opts would use (nel
 * values, host; the learned and scaled kinds need their network / scaling set) */
int ust_get_schur_preconditioner(ust_context *ctx, const ust_solve_opts *opts, double *pc);
int ust_get_solution_dev(ust_context *ctx, const double **u_nodal_dev, const double **p_dev);
/* ---- PETSc binary dumps (-matdumpdir / writeMatVec.c role, auglag-driver-DGTGD.c:101-113) ----
 * Big-endian PETSc binary AIJ matrices and vectors, the format of the reference's .matrix fixtures.  K and G
 * are formed from the element matrices of the current assembly (structural Q1/P0 pattern); free_only drops
 * the Dirichlet velocity dofs and numbers the rest in node-major order (the reference's equation numbers). */
#define UST_MAT_K 0
#define UST_MAT_G 1
#define UST_VEC_F 0 /* momentum right-hand side (lifted) */
#define UST_VEC_H 1 /* continuity right-hand side */
#define UST_VEC_U 2 /* nodal velocity solution (BC values included) */
#define UST_VEC_P 3 /* element pressure */
int ust_write_petsc_matrix(ust_context *ctx, int which, int free_only, const char *path);
int ust_write_petsc_vector(ust_context *ctx, int which, int free_only, const char *path);

/* ---- FeVariable checkpoint files, HDF5 (FeVariable_SaveToFile, StgFEM/Discretisation/src/FeVariable.c:2136-2389;
 * FeVariable_ReadFromFile, :2452-2730) ----
 * Root attributes "checkpoint file version" (2), "dimensions" and, for a field on the Cartesian linear mesh,
 * "mesh resolution" (elements per axis), 32-bit big-endian integers; dataset "/data" of doubles, one row per
 * mesh node in global order: the node's coordinates first when saved, then the components.  The HDF5 C library
 * (1.10 or later) is loaded at the first call: UST_HDF5_LIB names it, else libhdf5.so.103 / libhdf5.so on the
 * loader's path or under /opt/conda/lib; UST_EINVAL when none loads. */
#define UST_FIELD_VELOCITY 0    /* VelocityField: linear mesh, dim components (the solution, BC values included) */
#define UST_FIELD_PRESSURE 1    /* PressureField: constant mesh (one node per element, at its centre), 1 component */
#define UST_FIELD_TEMPERATURE 2 /* TemperatureField: linear mesh, 1 component */
int ust_save_field_hdf5(ust_context *ctx, int field, int save_coords, const char *path);
/* reads a field back with FeVariable_ReadFromFile's checks (known version, dimensions, the mesh resolution of
 * a linear-mesh field, node count; coordinates detected from the row width; version-1 rows skip one leading
 * column); the velocity and pressure become the context's solution (the initial guess of the next solve), the
 * temperature its temperature field */
int ust_load_field_hdf5(ust_context *ctx, int field, const char *path);
/* any checkpoint file, without a context: header (version 1 when the attribute is absent; res[3] zero where
 * "mesh resolution" is absent), rows x cols data (row-major), writing one; the last failure's text in
 * ust_hdf5_error() */
int ust_hdf5_field_info(const char *path, int *version, int *dim, int *res, long *rows, int *cols);
int ust_hdf5_field_read(const char *path, double *data, long rows, int cols);
int ust_hdf5_field_write(const char *path, int dim, const int *res, long rows, int cols, const double *data);
const char *ust_hdf5_error(void);

/* ---- Uzawa solve (Stokes_SLE_UzawaSolver, Stokes_SLE_UzawaSolver.c:407-983) ----
 * Preconditioned CG on the pressure, Schur complement G^T K^-1 G applied through one velocity solve per
 * iteration; starts from the context's current pressure (the reference's pSolnVec), as the reference does.
 * The velocity solves use the multigrid PCG of `vel` (precision, levels, smoother; its tolerances are set
 * here: `tolerance` for the first solve, 0.1 `tolerance` for all others, Stokes_SLE_UzawaSolver.c:566, 600).
 * The pressure null space is not removed (as in the reference).  Errors as the reference's firewalls: a zero
 * momentum force vector (UST_EINVAL), a non-finite step or residual, and more than max_iterations
 * iterations (UST_ENOTCONV). */
typedef struct {
    double tolerance;            /* "tolerance" (default 1e-5) */
    int max_iterations;          /* "maxIterations" (1000) */
    int min_iterations;          /* "minIterations" (1) */
    int use_absolute_tolerance;  /* "useAbsoluteTolerance" (0): 1 stops on ||r|| / sqrt(n_p), 0 on that over
                                    the same norm of the Uzawa right-hand side G^T K^-1 f - h */
    int preconditioner;          /* 1: the UzawaPreconditionerTerm matrix approxS ("Preconditioner"), 0: none */
} ust_uzawa_opts;

void ust_uzawa_opts_default(ust_uzawa_opts *o);
/* stats: outer_its = Uzawa iterations; momentum_residual = ||f - K u - G p|| / ||[f; h]|| as for ust_solve */
int ust_solve_uzawa(ust_context *ctx, const ust_uzawa_opts *uzawa, const ust_solve_opts *vel, ust_solve_stats *stats);

/* K u = b velocity solve alone (one A11 solve) on free dofs; b, u nodal */
int ust_velocity_solve(ust_context *ctx, const ust_solve_opts *opts, const double *b, double *u, int *its);

/* ---- nonlinear Picard loop (SystemLinearEquations_NonLinearExecute, SystemLinearEquations.c:750-934) ---- */
int ust_picard(ust_context *ctx, double gravity, double tol, int max_its, const ust_solve_opts *opts,
               int *its_done, double *final_change, ust_solve_stats *last_stats);

/* ---- learned Schur scaling (MLP 27 -> H -> H -> 1, BF16 weights, FP32 accumulation) ---- */
/* feature k of pressure cell e = log10 eta(neighbour k) - log10 eta(e) over the 3^dim neighbourhood
 * (k = dx+1 + 3(dy+1) + 9(dz+1)), eta the element mean viscosity recovered from the K weights; a
 * neighbour outside the box reads 3.0 */
int ust_mlp_init(ust_context *ctx, int hidden, unsigned long long seed);
int ust_mlp_set_params(ust_context *ctx, int hidden, const float *w, const float *b);
int ust_mlp_get_params(ust_context *ctx, float *w, float *b);
int ust_mlp_param_counts(int dim, int hidden, int *nw, int *nb);
/* features of the current assembly (nel * 3^dim) */
int ust_mlp_features(ust_context *ctx, float *feat);
/* target for training: log10 of the ratio between a probed Schur diagonal and approxS */
/* per-cell log10 factor on approxS for UST_SCHUR_SCALED (e.g. the probed target itself), nel values */
int ust_set_schur_scaling(ust_context *ctx, const float *log10_scale);
int ust_mlp_probe_target(ust_context *ctx, const ust_solve_opts *opts, int nprobe, unsigned long long seed,
                         float *target);
/* n samples (host arrays) -> epochs of Adam on the GPU; returns the final loss */
int ust_mlp_train(ust_context *ctx, int n, const float *feat, const float *target, int epochs, int batch,
                  float lr, unsigned long long seed, float *final_loss);
int ust_mlp_predict(ust_context *ctx, int n, const float *feat, float *out);
/* gradients of the loss for one batch (for checking against a CPU reference) */
int ust_mlp_gradients(ust_context *ctx, int n, const float *feat, const float *target, float *grad_w,
                      float *grad_b, float *loss);

/* ---- reference plugin API mirror (StiffnessMatrixTerm_AssembleElementFunction,
 * StiffnessMatrixTerm.h:52-58): bind a term object to a context, then the callbacks fill
 * elStiffMat (+=) from the batched assembly ---- */
int ust_stg_bind_term(void *term, ust_context *ctx, int kind);
#define UST_TERM_CONSTITUTIVE 1   /* ConstitutiveMatrixCartesian */
#define UST_TERM_GRADIENT 2       /* GradientStiffnessMatrixTerm */
#define UST_TERM_UZAWA_PC 3       /* UzawaPreconditionerTerm */
#define UST_TERM_BUOYANCY 4       /* BuoyancyForceTerm */
void ust_ConstitutiveMatrixCartesian_AssembleElement(void *constitutiveMatrix, void *stiffnessMatrix,
                                                     int lElement_I, void *sle, void *context,
                                                     double **elStiffMat);
void ust_GradientStiffnessMatrixTerm_AssembleElement(void *matrixTerm, void *stiffnessMatrix, int lElement_I,
                                                     void *sle, void *context, double **elStiffMat);
void ust_UzawaPreconditionerTerm_AssembleElement(void *matrixTerm, void *stiffnessMatrix, int lElement_I,
                                                 void *sle, void *context, double **elPreconditioner);
void ust_BuoyancyForceTerm_AssembleElement(void *forceTerm, void *forceVector, int lElement_I,
                                           double *elForceVec);

#ifdef __cplusplus
}
#endif
#endif
