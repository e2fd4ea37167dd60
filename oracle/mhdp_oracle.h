/*
 * CPU oracle for the MAC pressure projection — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's pressure projection
 * (sdfgeoff/mhd-simulator), used as the checker for the HIP library and as the
 * timed CPU baseline.  It is never linked into the product library.
 *
 * It follows, function by function:
 *   src/fluid/solver.wgsl:17-107     (MAC helpers, mobility, cached/uncached stencil, rhs)
 *   src/fluid/solver.wgsl:164-250    (guess, refresh, reductions, apply, update, search, correct)
 *   src/fluid/solver.wgsl:317-328    (independent physical residual check)
 *   src/fluid/pressure_subspace.wgsl (mean-free residual projection)
 *   src/fluid/multigrid.wgsl:1-144   (aggregation build, V(1,1) Jacobi cycle, coarse solve)
 *   src/periodic.wgsl:5-10           (unsigned Euclidean wrap)
 *   src/fluid/mod.rs:969-1115        (project(): staging, batches, stopping rule, checks)
 *   src/fluid/mod.rs:1447-1493       (anchors(): BFS components and gauge encoding)
 *   src/fluid/multigrid.rs:25-82     (hierarchy(), cycle_stages())
 *
 * Two arithmetic modes share one code body (core.inc):
 *   32: FP32 with the reference's operation order, including its two-level
 *       256-wide tree reductions, so per-step rounding follows the WGSL.
 *   64: FP64 on the same discrete system (FP32 face mobilities promoted, RHS in
 *       FP64 from the FP32 face velocities) — the ground truth for 1e-8 parity.
 *
 * Pinning (what each part is checked against) is listed in tests/test_oracle.py;
 * exact iteration counts and residual histories of the reference run are not
 * pinned by anything available offline ("parity unpinned" for those).
 *
 * Layout conventions (shared with include/mhdp.h):
 *   cells  x + nx*(y + ny*z)
 *   faces  three SoA planes; plane a has dims (n + e_a), index
 *          x + Px*(y + Py*z) with P = n + e_a.  Periodic duplicates at
 *          index n_a are ignored on input and written on output.
 */
#ifndef MHDP_ORACLE_H
#define MHDP_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t nx, ny, nz;
    uint32_t periodic; /* bit a set: axis a periodic */
    float h;           /* cell size (Params.physical.x) */
    int two_phase;     /* Params.surface.x > 0.5 */
    float rho_liquid;  /* Params.physical.w */
    float rho_air;     /* Params.surface.y */
} orc_grid;

typedef struct {
    float rel_tol;       /* FluidConfig.relative_tolerance */
    float abs_div_tol;   /* FluidConfig.absolute_divergence_tolerance */
    uint32_t max_iter;   /* FluidConfig.max_iterations */
    int precond;         /* 0 auto (MG iff cells >= 96^3), 1 Jacobi, 2 MG */
    int gauge;           /* 0 auto (mean-free iff one component), 2 pinned */
    int precision;       /* 32 or 64 */
    double tol64;        /* precision 64: relative tolerance (no 1e-7 floor); <=0 uses rel_tol */
    int threads;         /* OpenMP threads, <=0 = default */
} orc_opts;

typedef struct {
    uint32_t iterations;
    uint32_t levels;        /* MG levels incl. fine; 0 for Jacobi */
    uint32_t components;
    uint32_t fluid_cells;
    int meanfree;
    int multigrid;
    double rel_residual;    /* sqrt(rr/bb) as reported by convergence() */
    double seconds;
} orc_stats;

/* Status codes (mirroring the reference's error strings, src/fluid/mod.rs:974,1079,1097) */
enum { ORC_OK = 0, ORC_NOT_CONVERGED = 1, ORC_NO_PROGRESS = 2, ORC_NONFINITE = 3, ORC_INVALID = 4 };

uint64_t orc_face_plane_size(const orc_grid *g, int axis);

/* anchors(): mask in = 0/1 per cell, out = 0 | (1|root<<2) | (2|root<<2); returns components. */
uint32_t orc_anchors(uint32_t *mask, const uint32_t dims[3], uint32_t periodic);

/* hierarchy(): fills level dims (4 u32 each: nx,ny,nz,count) and returns the level count
 * (fine included), or -1 on the reference's error cases.  *coarse_cells = sum of coarse counts. */
int orc_hierarchy(const uint32_t dims[4], uint32_t *level_dims, int max_levels, uint64_t *coarse_cells);

/* pressure_coefficients: beta planes (3, packed x|y|z) and the cached diagonal per cell. */
void orc_coefficients(const orc_grid *g, const uint32_t *mask, const float *alpha, int meanfree,
                      float *beta, float *diag);
/* the uncached diagonal() for every fluid cell (0 elsewhere) */
void orc_diagonal_uncached(const orc_grid *g, const uint32_t *mask, const float *alpha, int meanfree,
                           float *diag);
/* rhs(): b = -h^2 div(u*) over open faces, 0 at pinned anchors (zero_anchor=1) */
void orc_rhs_f32(const orc_grid *g, const uint32_t *mask, int zero_anchor, const float *ustar, float *b);
void orc_rhs_f64(const orc_grid *g, const uint32_t *mask, int zero_anchor, const float *ustar, double *b);
/* apply(): cached 7-point operator (identity at pinned anchors) */
void orc_apply_f32(const orc_grid *g, const uint32_t *mask, int meanfree, const float *beta,
                   const float *v, float *out);
void orc_apply_f64(const orc_grid *g, const uint32_t *mask, int meanfree, const float *beta,
                   const double *v, double *out);
/* laplacian(): uncached physical operator on every fluid row */
void orc_laplacian_f32(const orc_grid *g, const uint32_t *mask, const float *alpha, const float *v, float *out);
void orc_laplacian_f64(const orc_grid *g, const uint32_t *mask, const float *alpha, const double *v, double *out);
/* FP64 ||b - L phi|| / ||b|| over every fluid row (physical operator, anchors included). */
double orc_true_residual(const orc_grid *g, const uint32_t *mask, const float *alpha,
                         const float *ustar, const double *phi);
/* correct(): u = u* - beta (phi_R - phi_L)/h on open faces, 0 elsewhere (all plane entries). */
void orc_correct_f64(const orc_grid *g, const uint32_t *mask, const float *alpha, const float *ustar,
                     const double *phi, float *u_out);

/* FluidState::project(false): encoded mask in, phi (potential) and projected faces out. */
int orc_project(const orc_grid *g, const uint32_t *mask, const float *alpha, const float *ustar,
                const orc_opts *o, double *phi_out, float *u_out, orc_stats *st);

/* ---- fields() + diagnostics() (solver.wgsl:256-291, mod.rs:1116-1146) ---- */
typedef struct {
    float dt;              /* Params.physical.y */
    float acceleration[3]; /* Params.acceleration.xyz */
    int mushy;             /* Params.flags.w != 0 */
    int meanfree;          /* mean_free_pressure() */
} orc_fields_params;
/* cells: 16 floats per cell (velocity.xyzw, pressure.xyzw, force.xyzw, reserved.xyzw), read for
 * the reserved lanes and overwritten; em: 16 floats per cell (force at 8..10) or NULL;
 * totals[10]: max speed, max |div|, sum div^2, sum kinetic energy, max |p|, sum div, max |force|,
 * non-finite cells, max |face velocity|, sum melt (FP64 sums, in cell order). */
void orc_fields(const orc_grid *g, const uint32_t *mask, const float *faces, const double *phi, const float *em,
                const orc_fields_params *p, float *cells, double *totals);

/* ---- momentum step either side of the projection: predict + viscous (solver.wgsl:127-163) ---- */
typedef struct {
    float dt;              /* Params.physical.y */
    float viscosity;       /* kinematic viscosity (Params.physical.z) */
    float acceleration[3]; /* Params.acceleration.xyz */
    float surface_tension; /* Params.surface.z (two-phase capillary force) */
    int advect;            /* Params.flags.y != 0: semi-Lagrangian RK2 advection */
} orc_step_params;
/* predict (solver.wgsl:127-140): faces (three planes) -> out (three planes); closed faces 0. */
void orc_predict(const orc_grid *g, const uint32_t *mask, const float *faces, const orc_step_params *p, float *out);
/* viscous (solver.wgsl:143-163), mushy factor off: pred (three planes) -> out; cells: 16 floats per
 * cell (reserved.z liquid fraction, reserved.w curvature; read only for two-phase grids); em: 16
 * floats per cell (force at 8..10) or NULL. */
void orc_viscous(const orc_grid *g, const uint32_t *mask, const float *pred, const float *cells, const float *em,
                 const orc_step_params *p, float *out);

/* ---- two-phase viscous stress diffusion: diffuse_surface_stress (src/fluid/mod.rs:1146-1234),
 * src/fluid/viscosity.wgsl, src/fluid/stabilized.{rs,wgsl}.  cells: FluidCell packets (reserved.z =
 * occupancy); nu: liquid kinematic viscosity (Params.physical.z); mu_air: ambient dynamic viscosity
 * (Params.surface.w); u: face planes. ---- */
float orc_stress_bound(const orc_grid *g, const uint32_t *mask, const float *cells, float nu, float mu_air);
void orc_stress_euler(const orc_grid *g, const uint32_t *mask, const float *cells, float nu, float mu_air,
                      float dt_sub, const float *u, float *out);
int orc_rkc_schedule(double dt, uint32_t euler_steps, float *rows, int max_rows);
int orc_surface_stress(const orc_grid *g, const uint32_t *mask, const float *cells, double dt, float nu, float mu_air,
                       uint32_t max_substeps, float *u, uint32_t *substeps, double *count);

/* ---- EM Jacobi-CG on the edge curl-curl operator (src/em/induction.wgsl, src/em/mod.rs:863-1010) ---- */
typedef struct {
    uint32_t nx, ny, nz; /* cells; edges of axis a: dims n + 1 except n_a on axis a, packed x|y|z */
    float h;             /* cell size (Params.physical.z) */
} orc_em_grid;
typedef struct {
    float dt;          /* Params.physical.x */
    float mu0;         /* Params.physical.y */
    float tol2;        /* relative tolerance squared, Params.physical.w */
    uint32_t max_iter; /* EmConfig.max_iterations */
    int magnetostatic; /* Params.control.y == 1 (initial solve): no mass term */
    /* nodal auxiliary-space correction (src/em/nodal.wgsl; the switch of EmState::advance_to,
     * src/em/mod.rs:887-960): 0 the reference's policy (an induction solve allowed more than 1024
     * iterations switches at a true-residual restart after 1024), -1 never, 1 switch after
     * nodal_after iterations (0: from the start; the reference's nodal_switch_after test hook) */
    int nodal;
    uint32_t nodal_after;
} orc_em_params;
typedef struct {
    uint32_t iterations;
    double rel_residual; /* sqrt(rr / bb) of the last reduction */
    double scale;        /* power-of-two residual scale picked by initial_reduce */
    double start_residual;       /* sqrt(rr / bb) of the first start (prediction_diagnostic[1]) */
    uint32_t prediction_dropped; /* the predicted start was rejected (prediction_diagnostic[2]) */
    uint32_t nodal;              /* the solve ran (or switched to) the nodal correction */
} orc_em_stats;
/* coefficients (induction.wgsl:82-99): materials one u32 per cell, props[2k], props[2k+1] = (sigma,
 * mu_r) of material k; vacuum = Params.control.z == 1.  Out per edge: sigma (mean of the four
 * cells), diag (sum of the four face reluctivities, 1 on boundary edges, + 1 per air endpoint),
 * flags (1 boundary, 2 lower air node, 4 upper air node, 8 unit-reluctivity air stencil); per face
 * (three planes, dims n + e_a): nu; check = (boundary conductor edges, min sigma > 0, max sigma,
 * conducting edges). */
void orc_em_coefficients(const orc_em_grid *g, const uint32_t *materials, const float *props, int vacuum,
                         float *sigma, float *diag, uint8_t *flags, float *nu, float check[4]);
/* apply_operator (induction.wgsl:160-170) */
void orc_em_apply_f32(const orc_em_grid *g, const orc_em_params *p, const float *sigma, const float *diag,
                      const uint8_t *flags, const float *nu, const float *v, float *q);
void orc_em_apply_f64(const orc_em_grid *g, const orc_em_params *p, const float *sigma, const float *diag,
                      const uint8_t *flags, const float *nu, const double *v, double *q);
/* One step's solve: as_now / as_prev the applied potential (source.x/.y), mot the motional source
 * (source[i + E].x, NULL = 0), a_prev the accepted potential; a_out = a_prev + increment.  With
 * a_before (the potential accepted one step earlier, NULL = none) and prediction > 0 (origin.w) the
 * increment starts from prediction * (a_prev - a_before) off the boundary, else from zero.  Returns 0
 * converged, 1 not converged, 3 non-finite. */
int orc_em_solve_f32(const orc_em_grid *g, const orc_em_params *p, const float *sigma, const float *diag,
                     const uint8_t *flags, const float *nu, const float *as_now, const float *as_prev,
                     const float *mot, const float *a_prev, const float *a_before, float prediction, float *a_out,
                     orc_em_stats *st);
int orc_em_solve_f64(const orc_em_grid *g, const orc_em_params *p, const float *sigma, const float *diag,
                     const uint8_t *flags, const float *nu, const float *as_now, const float *as_prev,
                     const float *mot, const float *a_prev, const float *a_before, float prediction, double *a_out,
                     orc_em_stats *st);

/* The nodal correction applied to r (nodal_build_initial + nodal_z, nodal.wgsl:9-30):
 * z = r / d + N(lower node) - N(upper node), N = node values (dims n + 1, 0 off the interior);
 * node_diag (may be NULL) receives diag(G^T A G) per node, nodal (may be NULL) the node values. */
void orc_em_nodal_f32(const orc_em_grid *g, const orc_em_params *p, const float *sigma, const float *diag,
                      const uint8_t *flags, const float *r, float *z, float *node_diag, float *nodal);
void orc_em_nodal_f64(const orc_em_grid *g, const orc_em_params *p, const float *sigma, const float *diag,
                      const uint8_t *flags, const double *r, double *z, double *node_diag, double *nodal);

/* ---- multigrid probes (mirror the reference's test-only entry points) ---- */
typedef struct orc_mg orc_mg;
orc_mg *orc_mg_new(const orc_grid *g, const uint32_t *mask, const float *alpha, int meanfree, int precision);
void orc_mg_free(orc_mg *m);
int orc_mg_levels(const orc_mg *m, uint32_t *dims4 /* [levels*4] */);
/* action of level l (l=0 fine, uses cached beta; l>=1 Galerkin coarse graph) */
void orc_mg_action(orc_mg *m, int level, const double *v, double *out);
/* one V-cycle z = M r on the fine level */
void orc_mg_vcycle(orc_mg *m, const double *r, double *z);
/* coarsest level: fused solve vs n separate weighted-Jacobi sweeps from zero */
void orc_mg_coarse_solve(orc_mg *m, const double *rhs, double *z);
void orc_mg_coarse_sweeps(orc_mg *m, const double *rhs, double *z, int sweeps);
/* coarse graph lanes of level l>=1: c (4 per cell: lower x,y,z edge, reaction) and diag */
void orc_mg_graph(const orc_mg *m, int level, double *coeff4, double *diag);

#ifdef __cplusplus
}
#endif
#endif
