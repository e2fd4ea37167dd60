// Host-callable launchers of the projection kernels (one HIP stream per context).
#pragma once
#include <vector>

#include "mhdp_device.h"

struct PhaseP {
    int two_phase;
    float rho_l, rho_air;
    int meanfree;
};

// Storage precision of the multigrid preconditioner.
enum Store : int { STORE_F32 = 0, STORE_BF16 = 1 };

// One multigrid level on the device.  Level 0 aliases the fine coefficient arrays.
struct LevelDev {
    Grid3 g;
    float *bx, *by, *bz, *react, *diag, *dinv;  // FP32 graph (level 0: fine mobilities)
    void *bx16, *by16, *bz16, *react16, *dinv16; // BF16 copies (mixed precision)
    void *r, *z, *t;                              // level vectors (storage type of the cycle)
    void* w = nullptr;                            // level 1: second-visit correction (W-cycle)
};

// setup.hip
// Returns true when the BF16 copies (all four pointers non-null) were written too.
// open8: the per-cell open-face bytes (the row-vector pass reads them instead of the mask)
bool launch_coefficients(hipStream_t s, const Grid3& g, const PhaseP& ph, const uint32_t* mask, const uint8_t* open8,
                         const float* alpha,
                         float* bx, float* by, float* bz, float* diag, float* dinv, void* bx16 = nullptr,
                         void* by16 = nullptr, void* bz16 = nullptr, void* dinv16 = nullptr);
// open8: the per-cell open-face bytes (the row-vector passes read them instead of the mask and
// the mobilities; null: the one-cell pass)
void launch_rhs_f32(hipStream_t s, const Grid3& g, float h, const uint32_t* mask, const uint8_t* open8, const float* bx,
                    const float* by, const float* bz, const float* ustar, float* b);
// launch_rhs_f64 and launch_outer_start(zero_x) in one pass (rows of four cells: nx % 4 == 0, and
// the open bytes; returns false and launches nothing otherwise)
bool launch_rhs_start(hipStream_t s, const Grid3& g, float h, const uint8_t* open8, const float* ustar, double* b,
                      bool meanfree, float* r32, float* x32, float* p0, float* p1, double* part, Scal* sc);
void launch_rhs_f64(hipStream_t s, const Grid3& g, float h, const uint32_t* mask, const uint8_t* open8, const float* bx,
                    const float* by, const float* bz, const float* ustar, double* b);
void launch_correct_f32(hipStream_t s, const Grid3& g, float h, const float* bx, const float* by, const float* bz,
                        const float* phi, const float* ustar, float* uout);
// also copies phi to phi_copy (when set) in the same pass; returns whether it did
bool launch_correct_f64(hipStream_t s, const Grid3& g, float h, const float* bx, const float* by, const float* bz,
                        const double* phi, const float* ustar, float* uout,
                        double* phi_copy = nullptr);
void launch_mg_build(hipStream_t s, const Grid3& cg, const Grid3& fg, bool fine, int meanfree, const uint32_t* mask,
                     const float* fdinv, const float* fbx, const float* fby, const float* fbz, const float* freact,
                     float* cbx, float* cby, float* cbz, float* creact, float* cdiag, float* cdinv,
                     void* const c16[5] = nullptr);  // c16: BF16 copies of bx, by, bz, react, dinv
void launch_to_bf16(hipStream_t s, uint32_t n, const float* src, void* dst);

// pcg.hip
void launch_refresh(hipStream_t s, const Grid3& g, const float* x, const float* b, const float* bx, const float* by,
                    const float* bz, const float* dinv, float* r, float* p, double* part, Scal* sc, int meanfree,
                    int check_only, int mode);
void launch_subspace(hipStream_t s, const Grid3& g, float* r, float* p, const float* dinv, double* part, Scal* sc,
                     int initial, int mode);
void launch_apply(hipStream_t s, const Grid3& g, const float* p, const float* bx, const float* by, const float* bz,
                  float* q, double* part, Scal* sc);
void launch_update(hipStream_t s, const Grid3& g, float* x, float* r, const float* p, const float* q,
                   const float* dinv, double* part, Scal* sc, int meanfree, int mode);
void launch_search(hipStream_t s, const Grid3& g, const float* r, const float* dinv, float* p, Scal* sc);
void launch_pdot(hipStream_t s, const Grid3& g, int store, const float* r, const void* z, const float* dinv, float* p,
                 double* part, Scal* sc, int initial, int mode);
void launch_psearch(hipStream_t s, const Grid3& g, int store, const void* z, const float* dinv, float* p, Scal* sc);
void launch_residual64(hipStream_t s, const Grid3& g, const double* x64, const double* b64, const float* bx,
                       const float* by, const float* bz, const float* dinv, float* r32, double* part, Scal* sc);
void launch_outer_start(hipStream_t s, const Grid3& g, const double* x64, const double* b64, const float* bx,
                        const float* by, const float* bz, const float* dinv, float* r32, float* x32, float* p0,
                        float* p1, double* part, Scal* sc, bool zero_x);
// launch_outer_start with the mobilities formed from the liquid fraction (null: single phase) and
// the open bytes (rows of four cells, nx % 4 == 0; returns false and launches nothing otherwise)
bool launch_outer_start_m(hipStream_t s, const Grid3& g, const double* x64, const double* b64, const float* alpha,
                          const uint8_t* open8, float rho_l, float rho_air, bool meanfree, float* r32, float* x32,
                          float* p0, float* p1, double* part, Scal* sc, bool zero_x);
void launch_outer_update(hipStream_t s, const Grid3& g, double* x64, const float* x, const float* p0, const float* p1,
                         const float* dinv, const Scal* sc, bool meanfree, bool assign);
void launch_xupdate64(hipStream_t s, const Grid3& g, double* x64, const float* e, const float* dinv);
void launch_fill_f32(hipStream_t s, uint32_t n, float* dst, float v);
void launch_fill_f64(hipStream_t s, uint32_t n, double* dst, double v);
void launch_apply_plain(hipStream_t s, const Grid3& g, const float* p, const float* bx, const float* by,
                        const float* bz, const float* dinv, int identity_inactive, float* q);
void launch_cast_f32_f64(hipStream_t s, uint32_t n, const float* src, double* dst);

// Optional event pairs around one kernel family (live per-kernel timing).
struct KernelTimer {
    int kernel = -1;                         // MHDP_K_* being timed, -1 off
    std::vector<hipEvent_t> ev;              // pool, two per launch
    size_t used = 0;
    void before(int k, hipStream_t s) {
        if (k != kernel) return;
        if (used + 2 > ev.size()) {
            for (int i = 0; i < 64; i++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return;
                ev.push_back(e);
            }
        }
        (void)hipEventRecord(ev[used], s);
    }
    void after(int k, hipStream_t s) {
        if (k != kernel || used + 2 > ev.size()) return;
        (void)hipEventRecord(ev[used + 1], s);
        used += 2;
    }
};

// mg.hip — one V-cycle z = M r on the fine level (cycle_stages, multigrid.rs:67-82)
void launch_vcycle(hipStream_t s, int store, int L, LevelDev* lv, const float* r0, const uint32_t* mask, Scal* sc,
                   int coarse_f64, int visits, void** zfinal, KernelTimer* timer = nullptr);

// mg.hip — the coarse part of the V-cycle: levels 1..L-1 given the level-1 residual in lv[1].r;
// returns the level-1 correction to prolong (lv[1].z if level 1 is the coarsest, else lv[1].t).
// visits: how many times the fine level visits level 1 (1: V-cycle; 2: W-cycle at level 1, the
// second visit solving for the residual the first leaves; needs lv[1].w and L >= 3)
// rb: red-black Gauss-Seidel smoothing on every level rb_level_ok admits (weighted Jacobi elsewhere)
void launch_vcycle_coarse(hipStream_t s, int store, int L, LevelDev* lv, Scal* sc, int coarse_f64, int visits,
                          void** z1out, bool rb = false);

// fused.hip — the fused fine-grid MG-PCG passes
bool fused_vec4(const Grid3& g);
void launch_fused_apply(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                        const float* bx, const float* by, const float* bz, const float* dinv, float* q, double* part,
                        Scal* sc, int meanfree);
void launch_fused_xflush(hipStream_t s, const Grid3& g, const float* p0, const float* p1, float* x, const float* dinv,
                         Scal* sc, int meanfree);
void launch_clear_pending(hipStream_t s, Scal* sc);
void launch_fused_update(hipStream_t s, const Grid3& g, float* r, const float* q, const float* dinv, double* part,
                         Scal* sc);
void launch_fused_restrict(hipStream_t s, int store, const Grid3& cg, const Grid3& fg, const float* r, const void* dinv,
                           const void* bx, const void* by, const void* bz, void* rc, Scal* sc);
void launch_fused_smooth(hipStream_t s, int store, const Grid3& g, const Grid3& cg, const float* r, const void* dinv,
                         const void* bx, const void* by, const void* bz, const void* zc, void* t, double* part,
                         Scal* sc, int mode);

// brick.hip — 4x2x2-cell-per-thread variants (nx % 4 == 0, even ny and nz)
bool brick_ok(const Grid3& g);
bool brick_enabled_env();
// dinv: the storage type's reciprocal diagonal (BF16 copy with STORE_BF16), tested against zero only
// launch_brick_apply with the mobilities formed in the pass from the liquid fraction (null: single
// phase) and the open-face bytes (bit 4: a component's anchor), equal to the stored ones bit for bit
void launch_brick_apply_m(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                          const float* alpha, const uint8_t* open8, float rho_l, float rho_air, float* q,
                          double* part, Scal* sc, int meanfree);
// the same pass on the z-marching pair-column layout (nx % 4 == 0, with MHDP_APPLY_MARCH=1); false
// when it does not apply (nothing launched)
bool launch_apply_march(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                        const float* alpha, const uint8_t* open8, float rho_l, float rho_air, float* q, double* part,
                        Scal* sc, int meanfree);
void launch_brick_apply(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                        const float* bx, const float* by, const float* bz, const void* dinv, float* q, double* part,
                        Scal* sc, int meanfree);
void launch_brick_restrict(hipStream_t s, int store, const Grid3& cg, const Grid3& g, const float* r, const void* dinv,
                           const void* bx, const void* by, const void* bz, void* rc, Scal* sc);
// r2: the residual's twin buffer (the kernel reads the one Scal.rsel names); use_next: the
// V-cycle's mean is Scal.mean_next (after the fused residual update) instead of Scal.mean.
void launch_brick_smooth(hipStream_t s, int store, const Grid3& g, const Grid3& cg, const float* r, const float* r2,
                         const void* dinv, const void* bx, const void* by, const void* bz, const void* zc, void* t,
                         double* part, Scal* sc, int mode, bool use_next);
// Residual update fused into the restriction: r' = (r - mean) - alpha q from the buffer Scal.rsel
// names into its twin (FIN_RESTRICT_UPD flips rsel).
void launch_brick_restrict_update(hipStream_t s, int store, const Grid3& cg, const Grid3& g, float* r, float* r2,
                                  const float* q, const void* dinv, const void* bx, const void* by, const void* bz,
                                  void* rc, double* part, Scal* sc);
void launch_brick_restrict_coarse(hipStream_t s, int store, const Grid3& cg, const Grid3& g, const void* r,
                                  const void* dinv, const void* bx, const void* by, const void* bz, const void* react,
                                  void* rc, Scal* sc);
void launch_brick_smooth_coarse(hipStream_t s, int store, const Grid3& g, const Grid3& cg, const void* r,
                                const void* dinv, const void* bx, const void* by, const void* bz, const void* react,
                                const void* zc, void* t, Scal* sc, const void* base = nullptr, float theta = 0.0f);
// r = r - A z on a coarse brick level, in place (the residual a visit leaves; W-cycle)
void launch_brick_coarse_residual(hipStream_t s, int store, const Grid3& g, const void* bx, const void* by,
                                  const void* bz, const void* react, const void* z, void* r, Scal* sc);

// march.hip — z-marching forms of the fine passes (nx % 4 == 0, even ny and nz), BF16 V-cycle
// storage only (with FP32 storage the three-plane window does not fit the register file).
bool march_ok(const Grid3& g);
void launch_march_smooth(hipStream_t s, const Grid3& g, const Grid3& cg, const float* r, const float* r2,
                         const void* dinv, const void* bx, const void* by, const void* bz, const void* zc, void* t,
                         double* part, Scal* sc, int mode, bool use_next);
void launch_march_restrict_update(hipStream_t s, const Grid3& cg, const Grid3& g, float* r, float* r2,
                                  const float* q, const void* dinv, const void* bx, const void* by, const void* bz,
                                  void* rc, double* part, Scal* sc);

// rbgs.hip — red-black Gauss-Seidel passes of the multigrid cycle (any level; fine: mean-free
// deferred mean and, for the restriction, the PCG residual update from r/r2 (Scal.rsel) into the
// other buffer when q is given).  rb_level_ok: every periodic axis of the level has even length.
bool rb_level_ok(const Grid3& g);
// The smallest red-black levels and the coarsest solve of a cycle in one launch of one block:
// rb_tail_level gives the first level the tail takes (L: none), launch_rb_tail runs levels T..L-1
// from the residual of level T and writes level T's correction, as the per-level passes do.
// MHDP_RB_TAIL=0 keeps the per-level passes.
int rb_tail_level(int L, const LevelDev* lv);
void launch_rb_tail(hipStream_t s, int store, bool coarse_f64, int L, LevelDev* lv, int T, Scal* sc);
void launch_rb_restrict(hipStream_t s, int store, bool coarse, const Grid3& cg, const Grid3& g, const void* r,
                        const void* r2, float* w0, float* w1, const float* q, const void* dinv, const void* bx,
                        const void* by, const void* bz, void* rc, double* part, Scal* sc);
void launch_rb_smooth(hipStream_t s, int store, bool coarse, const Grid3& g, const Grid3& cg, const void* r,
                      const void* r2, const void* dinv, const void* bx, const void* by, const void* bz, const void* zc,
                      void* t, const void* base, float theta, double* part, Scal* sc, int mode, bool use_next);

// setup.hip — warm start (pressure_warm / pressure_commit)
void launch_warm(hipStream_t s, const Grid3& g, float h, const uint32_t* mask, const double* prev, const float* bx,
                 const float* by, const float* bz, const float* ustar, float* uout);
void launch_commit(hipStream_t s, uint32_t n, const uint32_t* mask, const double* prev, double* phi);
// phi = the previous potential relative to each component's anchor (0 off the fluid)
void launch_old_potential(hipStream_t s, uint32_t n, const uint32_t* mask, const double* prev, double* phi);

// setup.hip — fields() + diagnostics(): FluidCell packets and the 10 totals (part: 10 doubles
// per block of grid_blocks(n))
constexpr int kFieldTotals = 10;
void launch_fields(hipStream_t s, const Grid3& g, float h, float rho_l, float rho_air, int two_phase, float dt,
                   const float acc[3], int mushy, int meanfree, const uint32_t* mask, const uint8_t* open8,
                   const float* faces,
                   const double* phi, const float* em, float* cells, double* part, double* totals);

// step.hip — predict and viscous (solver.wgsl:127-163).  open8: one byte per cell, bit a set when
// the cell's lower face on axis a is open (launch_open8 builds it from the encoded mask); uo:
// scratch of one float per face for predict's open-face copy.
void launch_open8(hipStream_t s, const Grid3& g, const uint32_t* mask, uint8_t* open8);
void launch_predict(hipStream_t s, const Grid3& g, float h, float dt, int advect, const uint8_t* open8,
                    const float* u, float* uo, float* out);
// two-phase: ak is scratch of one float2 per cell (the packets' liquid fraction and curvature)
void launch_viscous(hipStream_t s, const Grid3& g, float h, float rho_l, float rho_air, int two_phase, float dt,
                    float viscosity, const float acc[3], float sigma, const uint8_t* open8, const float* pred,
                    const float* cells, float2* ak, const float* em, float* out);

// step.hip — two-phase viscous stress diffusion (src/fluid/viscosity.wgsl, stabilized.wgsl): per cell
// viscosity and density, the stability bound (part: stress_bound_blocks floats), the stress tensor
// of a face field (cached: the RKC2 stages' form), one explicit substep, the RKC2 initial operator
// and one stage (row: mu, nu, 1 - mu - nu, operator, initial); the last three read the tensor
struct StressBufs {
    float h, nu, mu_air, rho_l, rho_air;
    const uint32_t* mask;
    const uint8_t* open8;  // bit 3: fluid cell
    const float *mu, *rho;
    float* tensor;  // stress_tensor_len floats
};
size_t stress_tensor_len(const Grid3& g);
void launch_stress_cells(hipStream_t s, const Grid3& g, const uint32_t* mask, const float* cells, float nu, float mu_air,
                         float rho_l, float rho_air, float* mu, float* rho);
uint32_t stress_bound_blocks(const Grid3& g);
void launch_stress_bound(hipStream_t s, const Grid3& g, const StressBufs& b, float* part, float* out);
void launch_stress_tensor(hipStream_t s, const Grid3& g, const StressBufs& b, int cached, const float* u);
void launch_stress_euler(hipStream_t s, const Grid3& g, const StressBufs& b, float dt, const float* u, float* out);
void launch_stress_initial(hipStream_t s, const Grid3& g, const StressBufs& b, const float* u, float* f0);
void launch_stress_stage(hipStream_t s, const Grid3& g, const StressBufs& b, const float row[5], const float* prev,
                         const float* older, const float* y0, const float* f0, float* out);

// em.hip — the EM step's Jacobi-CG on the edge curl-curl operator (src/em/induction.wgsl).
// Edges: per axis a dims n + 1 except n_a, packed x | y | z; faces: dims n + e_a, packed.
struct EmDims {
    uint32_t n[3];
    float h;
};
struct EmPhys {
    float dt, mu0;
    int magnetostatic;  // no mass term (the initial magnetostatic solve)
};
struct EmVecs {
    float *x, *r, *p, *q;                          // increment, residual, search direction, product
    const float *sigma, *diag, *nu;                // per edge, per edge, per face
    const uint8_t* flags;                          // per edge: 1 boundary, 2/4 lower/upper air, 8 air stencil
    const float *as_now, *as_prev, *mot, *a_prev;  // per edge (mot may be NULL)
    float *tface, *ndiv;                           // scratch: per face nu curl p, per node div p
};
// device scalars of the solve (FP64; status 0 running, 1 converged, 2 breakdown)
struct EmScal {
    double rz, rr, bb, l1, threshold, alpha, beta, scale, tol2;
    int status;
    uint32_t its;
    int rescale;
    int pad;
    double check[4];  // coefficient_check: boundary conductors, min sigma > 0, max sigma, conductors
};
uint32_t em_blocks(const EmDims& d);  // partial records (4 doubles each) of the edge kernels
void launch_em_coefficients(hipStream_t s, const EmDims& d, const uint32_t* materials, const float* props, int vacuum,
                            float* sigma, float* diag, uint8_t* flags, float* nu, double* part, void* scal);
// begin_step: x = prediction * (a_prev - a_before) off the boundary; prepare from x (guess 1) or
// from a zero increment (guess 0, x cleared)
void launch_em_begin(hipStream_t s, const EmDims& d, const uint8_t* flags, const float* a_prev, const float* a_before,
                     float prediction, float* x);
// nodal: the nodal auxiliary-space correction (src/em/nodal.wgsl) is on: prepare and restart start
// the search at its z, the iterations use it in place of r / d
void launch_em_prepare(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part, void* scal,
                       int guess, bool nodal = false);
void launch_em_iterations(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part, void* scal,
                          uint32_t count, bool nodal, KernelTimer* timer);
void launch_em_restart(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part, void* scal,
                       bool nodal = false);
void launch_em_probe_nodal(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part,
                           void* scal);
void launch_em_commit(hipStream_t s, const EmDims& d, const float* a_prev, const float* x, float* out);
void launch_em_probe_apply(hipStream_t s, const EmDims& d, const EmPhys& ph, const float* nu, const uint8_t* flags,
                           const float* sigma, const float* p, float* q, float* tface, float* ndiv);
