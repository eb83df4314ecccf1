/*
 * ust_kernels.h -- thin C ABI over the gfx950 HIP kernels of the PIC Stokes hot path.
 *
 * Plain device pointers and sizes, one HIP stream per call (passed as void*).  Every
 * launcher returns 0 on success or a negative UST_E* code; it never allocates.
 * The C host layer (ust.h) is built on these entry points only.
 *
 * Quadrature-weight representation (the MI355X-first form of the PIC element integrals):
 * on a rectilinear Q1 element with isotropic D, every integrand that
 * ConstitutiveMatrixCartesian_AssembleElement (ConstitutiveMatrixCartesian.c:277-410)
 * sums over the particles is a polynomial of degree <= 2 per reference axis.  The particle
 * sum is therefore reproduced exactly by the 3^dim Lagrange points xi in {-1,0,1}^dim with
 * weights  WK[q] = sum_p w_p detJ eta_p L_q(xi_p),  and
 *   K_e = sum_q WK[q] B(xi_q)^T D0 B(xi_q),    D0 = diag(2,..,2,1,..,1).
 * G_e (GradientStiffnessMatrixTerm.c:220-249) and f_e (BuoyancyForceTerm.c:435-448) are
 * multilinear: WG[b] = sum_p w_p detJ N_b(xi_p), WF[b] = sum_p w_p detJ rho_p N_b(xi_p).
 * WG and WF are SoA (W[b*nel + e]); WK is tile-blocked (ust_k_weights_len below).
 */
#ifndef UST_KERNELS_H
#define UST_KERNELS_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define UST_OK 0
#define UST_EINVAL (-1)
#define UST_EDEVICE (-2)
#define UST_ENOMEM (-3)
#define UST_ESTATE (-4)
#define UST_ENOTCONV (-5)

/* Rheology kinds (Underworld/Rheology component Type names).  Parameter slots p[]:
 *   MaterialViscosity  p0 eta0
 *   Arrhenius          p0 eta0, p1 activationEnergy, p2 activationVolume, p3 referenceTemperature
 *   ViscosityLimiter   p0 maxViscosity, p1 minViscosity
 *   VonMises           p0 cohesion, p1 minimumViscosity, p2 strainRateSoftening (0/1),
 *                      p3 cohesionAfterSoftening, p4 StrainWeakening (1: softened by the strain-weakening state)
 *   DruckerPrager      p0 cohesion, p1 frictionCoefficient, p2 minimumYieldStress, p3 minimumViscosity,
 *                      p4 strainRateSoftening (0/1), p5 cohesionAfterSoftening,
 *                      p6 frictionCoefficientAfterSoftening, p7 StrainWeakening (0/1) */
#define UST_RH_NONE 0
#define UST_RH_MATERIALVISCOSITY 1
#define UST_RH_ARRHENIUS 2
#define UST_RH_VISCOSITYLIMITER 3
#define UST_RH_VONMISES 4
#define UST_RH_DRUCKERPRAGER 5
#define UST_MAX_RHEOLOGIES 4
#define UST_MAX_MATERIALS 64

typedef struct {
    int type;
    double p[8];
} ust_rheology;

typedef struct {
    int nrh;
    ust_rheology rh[UST_MAX_RHEOLOGIES];
    /* BuoyancyForceTerm material extension (BuoyancyForceTerm.c:118-122): "density", "alpha",
     * "referenceTemperature", "beta", "referenceDepth" (all 0 by default) */
    double density, alpha, ref_temperature, beta, ref_depth;
} ust_material;

/* Per-particle state of the yield rheologies.  pfw: StrainWeakening's postFailureWeakening (StrainWeakening.h
 * particle extension; NULL: no strain weakening); the yield criterion of a rheology with the StrainWeakening
 * flag interpolates its parameters by ratio = min(max(pfw, 0) / softening_strain, 1) (StrainWeakening.c:
 * 484-503), and the assembly writes postFailureWeakeningIncrement to pfw_inc (StrainWeakening.c:448-482,
 * YieldRheology.c:224-226).  tensile: DruckerPrager's tensileFailure flag (DruckerPrager.c:312), or NULL. */
typedef struct {
    double *pfw;
    double *pfw_inc;
    unsigned char *tensile;
    double softening_strain; /* "softeningStrain" */
    double healing_rate;     /* "healingRate" */
} ust_yield_state;

/* Per-element population-control thresholds of PCDVC (PCDVC.c:239-285): upper_t / lower_t in percent of the
 * element volume, at most max_splits splits and max_deletions deletions per element, and whether interface
 * elements (particles of more than one material) may split or delete. */
typedef struct {
    double upper_t, lower_t;
    int max_deletions, max_splits;
    int split_in_interface_cells, delete_in_interface_cells;
} ust_pcdvc_opts;

/* BuoyancyForceTerm component keys (BuoyancyForceTerm.c:139-160): "VerticalAxis" (default 1),
 * "gravityDirection" (gHat, used when use_direction), "adjust", and the optional "HydrostaticTerm" whose
 * background density (HydrostaticTerm.c:267-330) is subtracted from every particle's density (its keys, and
 * hs_time for the context's current time, which moves the layer boundaries when hs_v != 0). */
typedef struct {
    int vertical_axis;
    int use_direction;
    double direction[3];
    int adjust;
    int hydrostatic;
    double hs_upper_density, hs_upper_alpha, hs_lower_density, hs_lower_alpha, hs_height, hs_material_boundary;
    double hs_T0, hs_linear_coefficient, hs_exponential_coefficient1, hs_exponential_coefficient2;
    double hs_v, hs_width, hs_time;
} ust_buoyancy;

/* Geometry of one (multigrid) level of the Cartesian mesh. */
typedef struct {
    int dim;
    int nel[3];
    double h[3];      /* element size per axis */
    double xmin[3];
    double ymax;      /* top of the box (Arrhenius depth) */
    double xmax[3];   /* box extent; vertices are xmin + i*(xmax-xmin)/nel as CartesianGenerator writes them */
} ust_grid;

/* Material shapes evaluated per particle (StgDomain/Shape role): later shapes win. */
#define UST_SHAPE_EVERYWHERE 0
#define UST_SHAPE_BELOWCOSINEPLANE 1 /* p: offset, amplitude, wavelength, phase (BelowCosinePlane.c:195-210) */
#define UST_SHAPE_SPHERE 2           /* p: cx, cy, cz, radius */
#define UST_SHAPE_BOX 3              /* p: minx, miny, minz, maxx, maxy, maxz */
typedef struct {
    int type;
    int material;
    double p[6];
} ust_shape;

/* --- swarm set-up --- */
/* RandomParticleLayout (RandomParticleLayout.c:195-213, Random.c:56-73): particle p of the cell-sorted
 * swarm takes MINSTD draws 3p+1..3p+3 of the stream seeded with seed^123459876 (jump-ahead, so every
 * particle is independent), xi by the Q1 inverse map, w = 2^dim / n_cell (ConstantWeights.c:180-189). */
int ust_k_random_layout(const ust_grid *g, const int *cell_off, int np, long seed, double *xi0, double *xi1,
                        double *xi2, double *w, void *stream);
/* SpaceFillerParticleLayout (SpaceFillerParticleLayout.c:218-240, GlobalParticleLayout.c:275-318): global
 * particle p (0..np-1) at min + (max - min) s_d(p+1) per axis, s_d the Sobol numbers of direction numbers
 * dirnum[d*32 + k] (k < 30); key[p] its element by Mesh_RegularAlgorithms_SearchElements (a point on an interior
 * face goes to the lower element), xi its local coordinates there.  The caller sorts by key. */
int ust_k_spacefiller_layout(const ust_grid *g, int np, const unsigned *dirnum, unsigned *key, double *xi0,
                             double *xi1, double *xi2, void *stream);
/* material of each particle: cell_material[e] (or 0) overridden by each containing shape in order */
int ust_k_assign_materials(const ust_grid *g, const int *cell_off, int np, const double *xi0, const double *xi1,
                           const double *xi2, const ust_shape *shapes_dev, int nshape,
                           const unsigned char *cell_material, unsigned char *mat, void *stream);
/* material of each particle from a nodal field (float, node-major) interpolated at the particle by the Q1
 * shape functions of its element: first_material + round((v - lo) / (hi - lo) * (nbins - 1)), clamped to the
 * nbins materials (a continuous property field expressed as RheologyMaterials; no reference counterpart: the
 * reference builds such set-ups from many Shape components) */
int ust_k_assign_materials_field(const ust_grid *g, const int *cell_off, int np, const double *xi0,
                                 const double *xi1, const double *xi2, const float *field_nodal, double lo, double hi,
                                 int nbins, int first_material, unsigned char *mat, void *stream);
/* coh[p] = 1 + amp * u, u uniform in (-1, 1) from a counter hash of (seed, p) */
int ust_k_noise_scale(float *coh, int np, float amp, unsigned long long seed, void *stream);

/* Length (in elements of the weight type) of a K-weight array of grid g.  The K weights are stored
 * tile-blocked.  3D: columns of 16x16 elements in lexicographic order, each column layer by layer,
 * W[((column*nz + layer)*27 + q)*256 + local]; 2D: 16x16 tiles, W[(tile*9 + q)*256 + local]; local
 * index lexicographic inside the 16x16 block, partial blocks padded with zeros. */
size_t ust_k_weights_len(const ust_grid *g);
/* tile-blocked position of weight q of element e (the host-side view of the same layout) */
size_t ust_k_weights_index(const ust_grid *g, int e, int q);

/* --- memory / stream plumbing (HIP runtime behind a C ABI) --- */
int ust_dev_set(int device);
int ust_dev_malloc(void **p, size_t bytes);
int ust_dev_free(void *p);
int ust_host_malloc_pinned(void **p, size_t bytes);
int ust_host_free_pinned(void *p);
int ust_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int ust_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int ust_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int ust_stream_create(void **stream);
int ust_stream_destroy(void *stream);
int ust_stream_sync(void *stream);
int ust_event_create(void **ev);
int ust_event_destroy(void *ev);
int ust_event_record(void *ev, void *stream);
int ust_event_elapsed_ms(void *ev_start, void *ev_stop, float *ms);
/* capture the launches issued on stream between begin and end into an instantiated graph (exec);
 * abort ends a capture that failed part-way and drops it */
int ust_graph_begin(void *stream);
int ust_graph_end(void *stream, void **exec);
int ust_graph_abort(void *stream);
int ust_graph_launch(void *exec, void *stream);
int ust_graph_destroy(void *exec);
const char *ust_dev_error_string(void);

/* --- assembly (particle -> quadrature weights, fused per-particle rheology) --- */
/* eta_out/yield_out (np) may be NULL.  temperature (nodal), velocity (nodal, interleaved dim),
 * pressure (per element), coh_scale (per particle) may be NULL when no rheology needs them.
 * WK (nq per element, tile-blocked), WG (nv*nel), WF (nv*nel): nq = 3^dim, nv = 2^dim.  ppc_hint: the
 * mean particle count per cell (sizes the lane segment of an element; any count is handled).  yield_state
 * may be NULL (no strain weakening, no tensile flags). */
int ust_k_particle_weights(const ust_grid *g, const int *cell_off, int ppc_hint, const double *xi0,
                           const double *xi1, const double *xi2, const double *w,
                           const unsigned char *mat, const float *coh_scale,
                           const ust_material *materials_dev, int nmat, const double *temperature,
                           const double *velocity, const double *pressure, int prev_solution,
                           const ust_yield_state *yield_state, const ust_buoyancy *buoyancy, double *eta_out,
                           unsigned char *yield_out, double *WK, double *WG, double *WF, void *stream);
/* PCDVC population control (PCDVC.c:1338-1432 2D, :880-990 3D, non-Inflow branch), in two passes around a
 * host-side scan of the counts.  plan: one thread per element reads the DVC region volumes w of its particles
 * and writes its new particle count and its new particle list (entries s: original particle s of the element;
 * s | 2^30: a split copy of s) at list[cell_off[e] + e * max_splits], capacity n_e + max_splits; flag[0] gets 1
 * where a changed element keeps fewer than 3 particles, flag[1] and flag[2] count the splits and deletions.
 * build: writes the new swarm from the lists (split copies at the parent's region centroid c0/c1/c2, the
 * parent's material, cohesion factor and two optional per-particle states). */
int ust_k_pcdvc_plan(const ust_grid *g, const int *cell_off, const double *w, const unsigned char *mat,
                     const ust_pcdvc_opts *opts, int *new_count, int *list, int *flag, void *stream);
int ust_k_pcdvc_build(int dim, int nel, int max_splits, const int *cell_off, const int *new_off, const int *list,
                      const double *xi0, const double *xi1, const double *xi2, const double *c0, const double *c1,
                      const double *c2, const unsigned char *mat, const float *coh, const double *state0,
                      const double *state1, double *o0, double *o1, double *o2, unsigned char *omat, float *ocoh,
                      double *ostate0, double *ostate1, void *stream);
/* StrainWeakening time step: pfw += dt * pfw_inc, then negative values to 0 (the TimeIntegrand update of
 * postFailureWeakening, whose time derivative is the stored increment, StrainWeakening.c:401-407, and
 * _StrainWeakening_MakeValuesPositive, :411-426) */
int ust_k_strain_weakening_step(double *pfw, const double *pfw_inc, int np, double dt, void *stream);
/* dst[s] = src[perm[s]] for s < n (per-particle state carried through a swarm re-sort) */
int ust_k_permute_f64(const unsigned *perm, const double *src, double *dst, int n, void *stream);
/* BuoyancyForceTerm "adjust" (BuoyancyForceTerm.c:362-375): WF of each element times 2^dim / (sum of its particle
 * weights), the sum recovered from WG (sum_b WG[b] = detJ sum_p w_p) */
int ust_k_adjust_force_weights(const ust_grid *g, const double *WG, double *WF, void *stream);

/* K_e for elements [first, first+count): Ke (count * nd * nd) row-major, FP64 MFMA. */
int ust_k_element_matrices(const ust_grid *g, const double *WK, int first, int count, double *Ke,
                           void *stream);
/* G_e (count*nd) and f_e (count*nd) from the vertex weights */
int ust_k_element_G(const ust_grid *g, const double *WG, int first, int count, double *Ge, void *stream);
int ust_k_element_f(const ust_grid *g, const double *WF, double gravity, int first, int count,
                    double *fe, void *stream);
/* the same along a direction: fe[a*dim + d] = gravity * dir[d] * WF[a] (dir = gHat, or -1 on the vertical
 * axis; ust_k_element_f is dir = (0, -1, 0)) */
int ust_k_element_f_dir(const ust_grid *g, const double *WF, double gravity, const double *dir, int first,
                        int count, double *fe, void *stream);
/* approxS_e = sum_v G_e[v]^2 / K_e[v][v] (UzawaPreconditionerTerm.c:200-212) */
int ust_k_uzawa_diag(const ust_grid *g, const double *WK, const double *WG, double *approxS, void *stream);

/* --- swarm advection and re-sort (SwarmAdvector.c:288-313, TimeIntegrand.c:221-330) --- */
/* Per particle: position from (cell, xi), Euler (order 1) or midpoint (order 2) step through the Q1 velocity
 * u (nodal, interleaved, BC values included), new element key (nel: left the box) and local coordinates in
 * it (n0..n2).  A corrector outside the box falls back to the first-order step when `fallback`, else sets
 * bit 0 of *flag and leaves the particle in place. */
int ust_k_advect(const ust_grid *g, const int *cell_off, int np, const double *xi0, const double *xi1,
                 const double *xi2, const double *u, double dt, int order, int fallback, unsigned *key, double *n0,
                 double *n1, double *n2, int *flag, void *stream);
/* stable sort of particle indices by key (perm), work of ust_k_sort_by_cell_work_bytes bytes */
size_t ust_k_sort_by_cell_work_bytes(int np, int nel);
int ust_k_sort_by_cell(const unsigned *key, unsigned *key_sorted, unsigned *perm, unsigned *perm_tmp, int np, int nel,
                       void *work, size_t work_bytes, void *stream);
/* cell_off[0..nel] from sorted keys (empty cells included); cell_off[nel] = particles kept */
int ust_k_cell_offsets(const unsigned *key_sorted, int np, int nel, int *cell_off, void *stream);
/* the re-sorted swarm: local coordinates, ConstantWeights 2^dim / count (ConstantWeights.c:187), material and
 * noise factor (optional) gathered through perm */
int ust_k_swarm_gather(int dim, int np, const unsigned *perm, const unsigned *key_sorted, const int *cell_off,
                       const double *n0, const double *n1, const double *n2, const unsigned char *mat,
                       const float *coh, double *o0, double *o1, double *o2, double *w, unsigned char *omat,
                       float *ocoh, void *stream);

/* --- DVC integration weights (PICellerator/Weights/src/DVCWeights.c:203-925) --- */
/* LDS bytes of one element's workgroup for sub-grid res (dim entries) and max_np particles per element */
size_t ust_k_dvc_lds_bytes(int dim, const int *res, int max_np);
/* Per element of a cell-sorted swarm: discrete Voronoi regions of its particles on the res sub-grid of
 * [-1,1]^dim, weight w = region volume, xi moved to the region centroid when set_centroids.  Bits of *flag:
 * 1 an element without particles, 2 an element with more than max_np, 4 the round limit (never reached when
 * ownership only passes to strictly nearer particles).  res entries 1..64, at most 4096 sub-cells, LDS of
 * ust_k_dvc_lds_bytes <= 64 KiB. */
int ust_k_dvc_weights(int dim, const int *res, int nel, const int *cell_off, int max_np, double *xi0, double *xi1,
                      double *xi2, double *w, int set_centroids, int *flag, void *stream);

/* --- operators.  mask: per dof (n_nodes*dim), 1 = Dirichlet.  y is overwritten. --- */
/* y = K u on the free dofs (input and output BC dofs treated as zero).  work: device workspace of
 * ust_k_apply_work_len(g) elements of the operator's type (per-tile node partial sums); one
 * workspace may be shared by all calls on one stream. */
size_t ust_k_apply_work_len(const ust_grid *g);
int ust_k_apply_K_f64(const ust_grid *g, const double *WK, const unsigned char *mask, const double *u,
                      double *y, double *work, void *stream);
int ust_k_apply_K_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *u,
                      float *y, float *work, void *stream);
/* fused epilogues of the same product (one pass, no intermediate K u vector):
 * residual: r = b - K u on the free dofs, r = 0 on the Dirichlet dofs (the multigrid residual before
 * restriction, PETScMGSolver.c:641-716);
 * cheby: one Jacobi-Chebyshev smoothing step in place, r = b - K x, d = c1 d + c2 invd.*r, x += d
 * (first: c1 = 0 and d is not read; d may then be NULL, and is not written, when no later step needs it --
 * so also for the prolongation passes below). */
int ust_k_apply_K_residual_f64(const ust_grid *g, const double *WK, const unsigned char *mask, const double *u,
                               const double *b, double *r, double *work, void *stream);
int ust_k_apply_K_residual_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *u,
                               const float *b, float *r, float *work, void *stream);
/* the same Chebyshev passes out of place: the step reads x and writes xo (distinct vectors; d is read and
 * written per dof, or left unwritten when NULL as above).  On small 3D grids (at most 65,536 elements: the coarser multigrid levels) the FP32 pass is
 * one halo-tile launch without a workspace gather (work is then not used). */
int ust_k_apply_K_cheby_oop_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *x,
                                float *xo, const float *b, const float *invd, float *d, float c1, float c2, int first,
                                float *work, void *stream);
int ust_k_apply_K_cheby_prolong_oop_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *x,
                                        float *xo, const float *b, const float *invd, float *d, float c2,
                                        const float *xc, float *work, void *stream);
/* the first Jacobi-Chebyshev step from a zero guess, x = c2 invd.*b, and the residual r = b - K x (0 on the
 * Dirichlet dofs) in one pass: a V-cycle level's pre-smoothing of degree 1 and its residual
 * (PETScMGSolver.c:641-716); the same x and r as ust_k_cheby_first_f32 followed by the residual pass */
int ust_k_apply_K_residual_first_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *b,
                                     const float *invd, float *x, float *r, float c2, float *work, void *stream);
int ust_k_apply_K_cheby_f64(const ust_grid *g, const double *WK, const unsigned char *mask, double *x,
                            const double *b, const double *invd, double *d, double c1, double c2, int first,
                            double *work, void *stream);
int ust_k_apply_K_cheby_f32(const ust_grid *g, const float *WK, const unsigned char *mask, float *x,
                            const float *b, const float *invd, float *d, float c1, float c2, int first,
                            float *work, void *stream);
/* 3D: the multigrid prolongation x += P xc (xc on the grid with half the elements per axis; Dirichlet dofs
 * of x untouched) followed by one first Chebyshev step (c1 = 0) on the corrected x, in one pass */
int ust_k_apply_K_cheby_prolong_f64(const ust_grid *g, const double *WK, const unsigned char *mask, double *x,
                                    const double *b, const double *invd, double *d, double c2, const double *xc,
                                    double *work, void *stream);
int ust_k_apply_K_cheby_prolong_f32(const ust_grid *g, const float *WK, const unsigned char *mask, float *x,
                                    const float *b, const float *invd, float *d, float c2, const float *xc,
                                    float *work, void *stream);
/* the FP32 operator with bfloat16 K weights (half the weight stream; the multigrid V-cycle's operator):
 * the same four passes as the _f32 entry points above */
int ust_k_apply_K_f32_bf16w(const ust_grid *g, const unsigned short *WK, const unsigned char *mask, const float *u,
                            float *y, float *work, void *stream);
int ust_k_apply_K_residual_f32_bf16w(const ust_grid *g, const unsigned short *WK, const unsigned char *mask,
                                     const float *u, const float *b, float *r, float *work, void *stream);
int ust_k_apply_K_cheby_f32_bf16w(const ust_grid *g, const unsigned short *WK, const unsigned char *mask, float *x,
                                  const float *b, const float *invd, float *d, float c1, float c2, int first,
                                  float *work, void *stream);
int ust_k_apply_K_cheby_prolong_f32_bf16w(const ust_grid *g, const unsigned short *WK, const unsigned char *mask,
                                          float *x, const float *b, const float *invd, float *d, float c2,
                                          const float *xc, float *work, void *stream);
/* y = K u as ust_k_apply_K_f64, without a caller workspace.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  The tile partials go to a
 * static device array of ust_k_apply_noalloc_max_work() doubles (32 MiB) that each GPU holds until the process
 * ends; a grid whose ust_k_apply_work_len exceeds it is rejected (UST_EINVAL, text in ust_dev_error_string).
 * Calls on one GPU share the array: calls that overlap (other streams, other threads) race.  Allocates nothing. */
size_t ust_k_apply_noalloc_max_work(void);
int ust_k_apply_K_f64_noalloc(const ust_grid *g, const double *WK, const unsigned char *mask, const double *u,
                              double *y, void *stream);
/* same operator, one element per thread with per-dof atomic adds (reference variant for checks) */
int ust_k_apply_K_ref_f64(const ust_grid *g, const double *WK, const unsigned char *mask, const double *u,
                          double *y, void *stream);
int ust_k_apply_K_ref_f32(const ust_grid *g, const float *WK, const unsigned char *mask, const float *u,
                          float *y, void *stream);
/* y = -K_{free,bc} u_bc  (BC lifting: input restricted to BC dofs, output to free dofs), accumulated */
int ust_k_lift_K(const ust_grid *g, const double *WK, const unsigned char *mask, const double *ubc,
                 double *y, void *stream);
/* y = G p (velocity dofs, free only); ust_k_apply_Gt mode (bc_only): 0 q = G^T u on the free dofs,
 * 1 q -= G^T u_bc over the BC dofs (continuity lifting), 2 q_e = sum over free dofs of |G_e| (row
 * sums of |D| for the infinity norm), 3 q_e = sum over free dofs of G_e^2 u (u = diag(K)^-1: the diagonal
 * of G^T diag(K)^-1 G, BSSCR's "gkgdiag" Schur approximation, preconditioner.c:268-310) */
int ust_k_apply_G(const ust_grid *g, const double *WG, const unsigned char *mask, const double *p,
                  double *y, void *stream);
int ust_k_apply_Gt(const ust_grid *g, const double *WG, const unsigned char *mask, const double *u,
                   double *q, int bc_only, void *stream);
/* y = sum over the elements of |G| per velocity dof (free dofs; the row sums whose maximum is ||G||_inf) */
int ust_k_abs_G_rowsum(const ust_grid *g, const double *WG, const unsigned char *mask, double *y, void *stream);
/* nodal force vector f (free dofs) from WF; accumulated into f */
int ust_k_force(const ust_grid *g, const double *WF, double gravity, const unsigned char *mask, double *f,
                void *stream);
/* the same along dir (3 doubles): f[n, d] += gravity * dir[d] * sum of WF around n, free dofs */
int ust_k_force_dir(const ust_grid *g, const double *WF, double gravity, const double *dir,
                    const unsigned char *mask, double *f, void *stream);
/* inverse diagonal of K on free dofs (1 on BC dofs) */
int ust_k_inv_diag(const ust_grid *g, const double *WK, const unsigned char *mask, double *invd64,
                   float *invd32, void *stream);

/* --- multigrid transfers (fine grid gf, coarse has nel/2) --- */
int ust_k_coarsen_weights(const ust_grid *gf, const double *WKf, double *WKc, void *stream);
/* restriction bc = P^T rf over the free fine dofs, bc = 0 on the Dirichlet coarse dofs (maskc); maskf may be
 * NULL when rf is already zero on the Dirichlet fine dofs (as the residual pass leaves it) */
int ust_k_restrict_f64(const ust_grid *gf, const unsigned char *maskf, const unsigned char *maskc,
                       const double *rf, double *bc, void *stream);
int ust_k_prolong_add_f64(const ust_grid *gf, const unsigned char *maskf, const double *xc, double *xf,
                          void *stream);
int ust_k_restrict_f32(const ust_grid *gf, const unsigned char *maskf, const unsigned char *maskc,
                       const float *rf, float *bc, void *stream);
/* restriction followed, in the same pass, by the coarse level's first Jacobi-Chebyshev step from a zero
 * guess: dc = xc = c2 invdc bc */
int ust_k_restrict_cheby_first_f64(const ust_grid *gf, const unsigned char *maskf, const unsigned char *maskc,
                                   const double *rf, double *bc, const double *invdc, double *dc, double *xc,
                                   double c2, void *stream);
int ust_k_restrict_cheby_first_f32(const ust_grid *gf, const unsigned char *maskf, const unsigned char *maskc,
                                   const float *rf, float *bc, const float *invdc, float *dc, float *xc, float c2,
                                   void *stream);
int ust_k_prolong_add_f32(const ust_grid *gf, const unsigned char *maskf, const float *xc, float *xf,
                          void *stream);

/* --- dense coarse-grid solve (FP64) --- */
/* A (n x n, n = n_nodes*dim) from the weights; BC dofs get identity rows/columns */
int ust_k_dense_assemble(const ust_grid *g, const double *WK, const unsigned char *mask, double *A,
                         void *stream);
/* in-place inverse of an SPD matrix by blocked Gauss-Jordan without pivoting (blocks of 32 pivots, three
 * launches each; n <= 4096); work: 32 n + 1024 doubles */
int ust_k_dense_spd_inverse(double *A, int n, double *work, void *stream);
/* y = Ainv x in FP64; the f32 variant converts at the boundary */
int ust_k_dense_matvec_f64(const double *Ainv, int n, const double *x, double *y, void *stream);
int ust_k_dense_matvec_f32(const double *Ainv, int n, const float *x, float *y, void *stream);

/* --- vector kernels (scalars that the device computes live in device memory) --- */
int ust_k_fill_f64(double *x, double v, size_t n, void *stream);
int ust_k_fill_f32(float *x, float v, size_t n, void *stream);
/* deterministic pseudo-random fill in (-1, 1) (power-iteration start vectors, probes) */
int ust_k_fill_hash_f64(double *x, size_t n, unsigned long long seed, void *stream);
/* Rademacher (+-1) fill */
int ust_k_fill_sign_f64(double *x, size_t n, unsigned long long seed, void *stream);
int ust_k_copy_f64_to_f32(const double *x, float *y, size_t n, void *stream);
int ust_k_copy_f32_to_f64(const float *x, double *y, size_t n, void *stream);
/* y = bfloat16(float(x)), round to nearest even (the K weights of the multigrid V-cycle) */
int ust_k_copy_f64_to_bf16(const double *x, unsigned short *y, size_t n, void *stream);
int ust_k_axpy_f64(double a, const double *x, double *y, size_t n, void *stream);     /* y += a x */
int ust_k_axpby_f64(double a, const double *x, double b, double *y, size_t n, void *stream);
int ust_k_add_f32_to_f64(const float *x, double *y, size_t n, void *stream);        /* y += x */
int ust_k_sub_f64(const double *a, const double *b, double *y, size_t n, void *stream); /* y = a - b */
int ust_k_scale_mul_f64(const double *a, const double *b, double *y, size_t n, void *stream); /* y = a .* b */
int ust_k_div_f64(const double *a, const double *b, double *y, size_t n, void *stream); /* y = a ./ b */
/* deterministic two-stage dot products: out[0] = x.y ; partial must hold 1024 doubles */
int ust_k_dot_f64(const double *x, const double *y, size_t n, double *partial, double *out, void *stream);
int ust_k_dot_f32(const float *x, const float *y, size_t n, double *partial, double *out, void *stream);
/* max |x_i| -> out[0] */
int ust_k_absmax_f64(const double *x, size_t n, double *partial, double *out, void *stream);
/* sum(x) -> out[0] */
int ust_k_sum_f64(const double *x, size_t n, double *partial, double *out, void *stream);
/* x -= s[0]/n  (constant-pressure null space projection) */
/* x[i] = 0 where mask[i] (restrict a nodal vector to the free dofs) */
int ust_k_mask_zero_f64(double *x, const unsigned char *mask, size_t n, void *stream);
int ust_k_remove_mean_f64(double *x, size_t n, const double *sum, void *stream);

/* CG building blocks with device scalars: sc = {rz, pq, rz_old, ...} */
/* p = z + (rz_new/rz_old) p  (first: p = z) */
int ust_k_cg_update_p_f64(const double *z, double *p, const double *rz_new, const double *rz_old, int first,
                          size_t n, void *stream);
int ust_k_cg_update_p_f32(const float *z, float *p, const double *rz_new, const double *rz_old, int first,
                          size_t n, void *stream);
/* x += a p ; r -= a q with a = num/den */
int ust_k_cg_update_xr_f64(double *x, double *r, const double *p, const double *q, const double *num,
                           const double *den, size_t n, void *stream);
int ust_k_cg_update_xr_f32(float *x, float *r, const float *p, const float *q, const double *num,
                           const double *den, size_t n, void *stream);
/* Device-controlled CG (no host round trip per iteration).  ctl[4] on the device: ctl[0] = stopping
 * threshold rtol2 * b.b, ctl[1] = state (0 running, 1 converged, 2 NaN), ctl[2] = iterations, ctl[3] =
 * beta.  start: ctl from bb = b.b and rz_old = 1 (with p = 0 the first update gives p = z).  beta:
 * ctl[3] = rz/rz_old, rz_old = rz.  The guarded updates and check do nothing once ctl[1] != 0, so
 * iterations launched past convergence leave x and r unchanged (KSPSolve's loop with the rnorm test
 * of KSPConvergedDefault, evaluated on the device). */
int ust_k_cg_start(const double *bb, double rtol2, double *ctl, double *rz_old, void *stream);
int ust_k_cg_beta(const double *rz, double *rz_old, double *ctl, void *stream);
int ust_k_cg_p_guarded_f64(const double *z, double *p, const double *ctl, size_t n, void *stream);
int ust_k_cg_xr_guarded_f64(double *x, double *r, const double *p, const double *q, const double *rz,
                            const double *pq, const double *ctl, size_t n, void *stream);
int ust_k_cg_check(const double *rr, double *ctl, void *stream);
/* r . z (two-stage, into rz) fused with the beta update of ust_k_cg_beta; z in FP32 or FP64 */
int ust_k_dot_mixed_beta(const double *x, const float *y, size_t n, double *partial, double *rz, double *rz_old,
                         double *ctl, void *stream);
int ust_k_dot_beta_f64(const double *x, const double *y, size_t n, double *partial, double *rz, double *rz_old,
                       double *ctl, void *stream);
/* fused forms for the FP64 PCG around the FP32 V-cycle: r . z with z in FP32; p = z + beta p with z in FP32;
 * x += a p, r -= a q together with r . r (two-stage, into rr) and the convergence check of ust_k_cg_check */
int ust_k_dot_mixed(const double *x, const float *y, size_t n, double *partial, double *out, void *stream);
int ust_k_cg_p_guarded_zf(const float *z, double *p, const double *ctl, size_t n, void *stream);
int ust_k_cg_xr_rr_guarded_f64(double *x, double *r, const double *p, const double *q, const double *rz,
                               const double *pq, double *ctl, size_t n, double *partial, double *rr,
                               void *stream);
/* The FP32 velocity PCG of iterative refinement: the same steps on FP32 vectors (dots and r . r accumulated in
 * FP64 against the same control block), the scaled rounding of the FP64 residual and the FP64 correction. */
int ust_k_cg_p_guarded_f32(const float *z, float *p, const double *ctl, size_t n, void *stream);
int ust_k_cg_xr_rr_guarded_f32(float *x, float *r, const float *p, const float *q, const double *rz, const double *pq,
                               double *ctl, size_t n, double *partial, double *rr, void *stream);
int ust_k_dot_beta_f32(const float *x, const float *y, size_t n, double *partial, double *rz, double *rz_old,
                       double *ctl, void *stream);
int ust_k_copy_f64_to_f32_scaled(const double *x, float *y, double s, size_t n, void *stream);
int ust_k_axpy_f32_to_f64(double s, const float *x, double *y, size_t n, void *stream);

/* Chebyshev smoother step (Jacobi-preconditioned).  r = b - y (y = A x given), then
 * d = c1*d + c2*invd.*r ; x += d  (c1 = 0 on the first step) */
int ust_k_cheby_step_f64(const double *b, const double *Ax, const double *invd, double *d, double *x,
                         double c1, double c2, size_t n, void *stream);
int ust_k_cheby_step_f32(const float *b, const float *Ax, const float *invd, float *d, float *x, float c1,
                         float c2, size_t n, void *stream);
/* first Chebyshev step from x = 0: d = c2 invd.*b, x = d */
int ust_k_cheby_first_f64(const double *b, const double *invd, double *d, double *x, double c2, size_t n,
                          void *stream);
int ust_k_cheby_first_f32(const float *b, const float *invd, float *d, float *x, float c2, size_t n, void *stream);
/* the same first step on an FP64 right-hand side, which is also rounded into b32 for the rest of the V-cycle */
int ust_k_cheby_first_from_f64(const double *b, float *b32, const float *invd, float *d, float *x, float c2,
                               size_t n, void *stream);
/* r = b - Ax */
int ust_k_residual_f64(const double *b, const double *Ax, double *r, size_t n, void *stream);
int ust_k_residual_f32(const float *b, const float *Ax, float *r, size_t n, void *stream);

/* --- learned Schur-preconditioner scaling (MLP 27 -> H -> H -> 1, BF16 weights, FP32 accumulation) --- */
/* feature: log10 of the element viscosity average over the 3^dim neighbourhood (clamped at the box) */
int ust_k_mlp_features(const ust_grid *g, const double *WK, float *feat, void *stream);
/* out[e] = MLP(feat[e]) ; weights packed (see ust_mlp_layout) */
int ust_k_mlp_forward(const float *feat, int n, int nin, int hid, const unsigned short *wbf16, const float *bias,
                      float *out, float *act1, float *act2, void *stream);
/* gradients of 0.5*sum((out-target)^2)/n w.r.t. weights and biases (FP32, deterministic block
 * partials; act1 is a workspace of >= 256*(nw+nb) floats, act2 of >= 256 floats) */
int ust_k_mlp_backward(const float *feat, const float *target, int n, int nin, int hid,
                       const unsigned short *wbf16, const float *bias, const float *act1, const float *act2,
                       const float *out, float *grad_w, float *grad_b, float *loss, void *stream);
int ust_k_mlp_adam(float *master_w, unsigned short *wbf16, float *bias, const float *grad_w, const float *grad_b,
                   float *m_w, float *v_w, float *m_b, float *v_b, int nw, int nb, float lr, float beta1,
                   float beta2, float eps, int step, void *stream);
/* pc[e] = approxS[e] * 10^clip(out[e], -4, 4) */
int ust_k_apply_scaling(const double *approxS, const float *out, double *pc, int n, void *stream);
int ust_k_mlp_to_bf16(const float *x, unsigned short *y, int n, void *stream);
/* probing of the Schur diagonal with a 3^dim colouring of the pressure cells */
int ust_k_color_indicator(const ust_grid *g, int color, double *v, void *stream);
int ust_k_color_gather(const ust_grid *g, int color, const double *q, double *d, void *stream);
/* t[e] = log10(d[e] / approxS[e]) */
int ust_k_log_ratio(const double *d, const double *approxS, float *t, int n, void *stream);

#ifdef __cplusplus
}
#endif
#endif
