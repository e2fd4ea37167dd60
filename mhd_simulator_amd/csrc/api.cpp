// C ABI of libmhdp.so: the pressure projection context and the host-side orchestration
// of FluidState::project (src/fluid/mod.rs:969-1115) over the HIP kernels.
#include "mhdp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "mhdp_kernels.h"
#include "launch.h"

void launch_from_bf16(hipStream_t s, uint32_t n, const void* src, float* dst);
void launch_components(hipStream_t s, const Grid3& g, const uint8_t* fluid, uint32_t* mask, uint32_t* counts);

namespace {

constexpr uint32_t kMaxPartials = 4096;
constexpr uint32_t kMgMinCells = 96u * 96u * 96u;  // multigrid.rs:4
// Visits of level 1 per cycle.  Two keep the damped sum t1 + 0.8 t2 positive definite with the
// margin derived for coarse_visits_t; a third visit would start from the second's undamped
// correction and shrink that margin, so the option stops at two.
constexpr int kMaxVisits = 2;
// Status of a solve whose multigrid PCG broke down (r.z or p.Ap <= 0): internal, never returned.
constexpr int kBreakdown = 100;
// Iterations without a new least inner residual after which an inner solve with a tuned cycle is
// taken as broken down.  A cycle that is not positive definite can keep r.z and p.Ap positive
// while the residual wanders without converging; a converging multigrid PCG (about a decade per
// iteration) sets a new least residual every few iterations.  The reference's cycle is never cut.
constexpr uint32_t kStallIterations = 100;

// Multigrid cycle of one solve: smoother weight, coarse-correction scale, level-1 visits.
struct Cycle {
    double omega, pscale;
    int visits;
    bool rb = false;  // red-black Gauss-Seidel (else weighted Jacobi with weight omega)
};

struct CoarseArrays {
    float *bx = nullptr, *by = nullptr, *bz = nullptr, *react = nullptr, *diag = nullptr, *dinv = nullptr;
    void *bx16 = nullptr, *by16 = nullptr, *bz16 = nullptr, *react16 = nullptr, *dinv16 = nullptr;
    void *r = nullptr, *z = nullptr, *t = nullptr;
    void* w = nullptr;  // level 1 only: the second coarse visit's correction (W-cycle)
};

}  // namespace

struct mhdp_ctx {
    mhdp_grid grid{};
    mhdp_phase phase{};
    Grid3 g{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    bool labels_on_host = false;  // the last mhdp_set_cells labelled on the host
    uint8_t* fluid8 = nullptr;    // fluid flags as given (device staging for the labelling)
    uint32_t* cc_counts = nullptr;
    uint32_t components = 0, fluid_cells = 0, anchor0 = 0;
    bool cells = false;
    bool has_alpha = false;
    // fine-level device arrays
    uint32_t* mask = nullptr;
    float* alpha = nullptr;
    float *bx = nullptr, *by = nullptr, *bz = nullptr, *diag = nullptr, *dinv = nullptr;
    float *b = nullptr, *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
    float* p2 = nullptr;  // second search-direction buffer of the fused path
    float* r2 = nullptr;  // twin of r for the brick path's fused residual update
    void *z = nullptr, *t = nullptr;
    double *x64 = nullptr, *b64 = nullptr;
    double* phi_prev = nullptr;  // potential of the last successful projection (warm start)
    bool prev_valid = false;
    float* uwarm = nullptr;      // u* corrected by the previous potential (warm start)
    double* fields_part = nullptr;  // fields(): per-block totals, then the 10 folded totals
    double* fields_h = nullptr;     // pinned host copy of the 10 folded totals
    float* faces_open = nullptr;    // predict(): open-face copy of its input (one float per face)
    uint8_t* open8 = nullptr;       // step kernels: per cell, bit a = lower face on axis a open
    float *stress_mu = nullptr, *stress_rho = nullptr, *stress_part = nullptr;  // viscous stress diffusion
    float *faces_b = nullptr, *faces_y0 = nullptr, *faces_f0 = nullptr;        // its face buffers
    float* stress_tensor = nullptr;                                            // and its stress tensor
    float* cell_pair = nullptr;  // viscous (two-phase): liquid fraction and curvature per cell
    bool open8_fresh = false;       // open8 matches the current mask
    void *bx16 = nullptr, *by16 = nullptr, *bz16 = nullptr, *dinv16 = nullptr;
    // multigrid
    std::vector<CoarseArrays> coarse;  // index l-1 for level l
    std::vector<Grid3> levels;         // level grids, [0] = fine
    bool mg_alloc = false, mg16_alloc = false;
    bool fine16_fresh = false;  // BF16 copies of the fine coefficients match the FP32 ones
    // Optional multigrid (src/fluid/mod.rs:1031-1043): its storage is allocated the first time a
    // nonzero residual asks for it; a refused allocation is final for the context, whose
    // projections then use Jacobi (a fitting Jacobi solve never fails for the optional part).
    bool mg_refused = false;
    uint64_t mg_bytes = 0;      // device bytes of the multigrid storage
    int mg_fail_after = -1;     // test hook: the next multigrid allocation fails after this many arrays
    bool in_mg_alloc = false;   // alloc_mg is allocating (the hook counts only its arrays)
    Cycle test_cycle{0.0, 0.0, 0};  // test hook: the next projection's first cycle (visits 0: none)
    // scalars
    Scal* sc = nullptr;
    Scal* sc_h = nullptr;
    double* part = nullptr;
    // staging for the host-pointer entry point
    float *ustar_stage = nullptr, *uout_stage = nullptr;
    double* phi_stage = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t bytes = 0;
    bool fused_enabled = true;  // MHDP_UNFUSED=1 selects the one-kernel-per-stage path
    bool brick_enabled = true;  // MHDP_NO_BRICK=1 keeps the 4-cell-group fused kernels
    bool march_enabled = true;  // z-marching fine passes where the grid allows (MHDP_NO_MARCH=1: bricks)
    int coarse_visits = 1;      // level-1 visits per cycle (mhdp_opts.coarse_visits)
    bool rb = false;            // red-black Gauss-Seidel smoothing (mhdp_opts.smoother)
    KernelTimer timer;
    double timer_ms = 0.0;
    uint32_t timer_launches = 0;
};

namespace {

int fail(mhdp_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIP_OK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return fail(c, MHDP_DEVICE, std::string("HIP error: ") + \
                                                              hipGetErrorString(e_) + " at " #call); \
    } while (0)

int dalloc(mhdp_ctx* c, void** p, size_t bytes) {
    if (*p) return MHDP_OK;
    if (c->in_mg_alloc && c->mg_fail_after >= 0) {  // test hook: refuse as an exhausted device would
        if (c->mg_fail_after == 0) {
            c->mg_fail_after = -1;
            return fail(c, MHDP_NO_MEMORY, "Optional pressure preconditioner allocation refused");
        }
        c->mg_fail_after--;
    }
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, MHDP_NO_MEMORY, "Fluid allocation exceeds the available device memory");
    }
    c->bytes += bytes;
    return MHDP_OK;
}
void dfree(mhdp_ctx* c, void* p) {
    if (p) (void)hipFree(p);
    (void)c;
}

uint64_t face_count(const mhdp_grid& g) {
    return (uint64_t)(g.nx + 1) * g.ny * g.nz + (uint64_t)g.nx * (g.ny + 1) * g.nz +
           (uint64_t)g.nx * g.ny * (g.nz + 1);
}

// anchors(), src/fluid/mod.rs:1447-1493: BFS components over the fluid mask with periodic
// wrap; every cell of a component is labelled with its root (lowest flat index, since the
// outer scan runs in index order); the root is the anchor (2|root<<2), the others members
// (1|root<<2).
uint32_t label_components(std::vector<uint32_t>& mask, const uint32_t d[3], uint32_t periodic) {
    const uint64_t n = (uint64_t)d[0] * d[1] * d[2];
    std::vector<uint32_t> roots(n, 0);
    std::vector<uint32_t> queue(n > 0 ? n : 1);
    uint32_t components = 0;
    for (uint64_t root = 0; root < n; root++) {
        if (mask[root] != 1u) continue;
        components++;
        mask[root] = 4u;
        roots[root] = (uint32_t)root;
        uint64_t head = 0, tail = 0;
        queue[tail++] = (uint32_t)root;
        while (head < tail) {
            const uint32_t i = queue[head++];
            const int64_t p[3] = {(int64_t)(i % d[0]), (int64_t)((i / d[0]) % d[1]), (int64_t)(i / ((uint64_t)d[0] * d[1]))};
            for (int axis = 0; axis < 3; axis++) {
                for (int dir = -1; dir <= 1; dir += 2) {
                    int64_t q[3] = {p[0], p[1], p[2]};
                    int64_t v = p[axis] + dir;
                    if (periodic & (1u << axis)) {
                        const int64_t m = d[axis];
                        v = ((v % m) + m) % m;
                    } else if (v < 0 || v >= (int64_t)d[axis]) {
                        continue;
                    }
                    q[axis] = v;
                    const uint64_t j = (uint64_t)q[0] + (uint64_t)d[0] * ((uint64_t)q[1] + (uint64_t)d[1] * (uint64_t)q[2]);
                    if (mask[j] == 1u) {
                        mask[j] = 3u;
                        roots[j] = (uint32_t)root;
                        queue[tail++] = (uint32_t)j;
                    }
                }
            }
        }
    }
    for (uint64_t i = 0; i < n; i++) {
        if (mask[i] == 3u) mask[i] = 1u | (roots[i] << 2);
        else if (mask[i] == 4u) mask[i] = 2u | (roots[i] << 2);
    }
    return components;
}

// hierarchy(), multigrid.rs:25-65: ceil-half until every dimension is <= 4.
std::vector<Grid3> hierarchy(const Grid3& fine) {
    std::vector<Grid3> out{fine};
    while (out.back().nx > 4 || out.back().ny > 4 || out.back().nz > 4) {
        Grid3 c = out.back();
        c.nx = (c.nx + 1) / 2;
        c.ny = (c.ny + 1) / 2;
        c.nz = (c.nz + 1) / 2;
        c.n = (uint32_t)c.nx * (uint32_t)c.ny * (uint32_t)c.nz;
        out.push_back(c);
    }
    return out;
}

int validate_opts(mhdp_ctx* c, const mhdp_opts* o) {
    if (!o) return fail(c, MHDP_INVALID, "Missing projection options");
    if (!(o->rel_tol >= 1e-7f && o->rel_tol <= 1e-2f) || o->max_iter < 1 || o->max_iter > 20000 ||
        !(o->abs_div_tol >= 0.0f) || !std::isfinite(o->abs_div_tol))
        return fail(c, MHDP_INVALID, "Invalid fluid timestep, initial condition or solver limits/tolerances");
    if (o->precond < 0 || o->precond > 2 || (o->precision != MHDP_FP32 && o->precision != MHDP_MIXED) ||
        (o->gauge != MHDP_GAUGE_AUTO && o->gauge != MHDP_GAUGE_PINNED))
        return fail(c, MHDP_INVALID, "Invalid pressure solver selection");
    if (!(o->smooth_weight >= 0.0f && o->smooth_weight < 1.0f) || !(o->coarse_scale >= 0.0f && o->coarse_scale < 2.0f))
        return fail(c, MHDP_INVALID, "Invalid V-cycle smoother weight or coarse-correction scale");
    if (o->coarse_visits < 0 || o->coarse_visits > kMaxVisits)
        return fail(c, MHDP_INVALID, "Invalid multigrid coarse visits");
    if (o->smoother < MHDP_SMOOTH_AUTO || o->smoother > MHDP_SMOOTH_RED_BLACK)
        return fail(c, MHDP_INVALID, "Invalid multigrid smoother");
    if (o->precision == MHDP_MIXED &&
        (!(o->refine_tol > 0.0) || !(o->refine_tol < 1.0) || o->inner_tol < 0.0f || o->inner_tol >= 1.0f))
        return fail(c, MHDP_INVALID, "Invalid refinement tolerance");
    return MHDP_OK;
}

int alloc_mg(mhdp_ctx* c, bool f16) {
    int rc;
    struct Account {  // every byte this call allocates is multigrid storage
        mhdp_ctx* c;
        uint64_t before;
        ~Account() {
            c->mg_bytes += c->bytes - before;
            c->in_mg_alloc = false;
        }
    } account{c, c->bytes};
    c->in_mg_alloc = true;
    if (!c->mg_alloc) {
        c->levels = hierarchy(c->g);
        c->coarse.assign(c->levels.size() - 1, CoarseArrays{});
        for (size_t l = 1; l < c->levels.size(); l++) {
            CoarseArrays& a = c->coarse[l - 1];
            const size_t n = c->levels[l].n;
            const size_t fb = n * sizeof(float);
            if ((rc = dalloc(c, (void**)&a.bx, fb)) || (rc = dalloc(c, (void**)&a.by, fb)) ||
                (rc = dalloc(c, (void**)&a.bz, fb)) || (rc = dalloc(c, (void**)&a.react, fb)) ||
                (rc = dalloc(c, (void**)&a.diag, fb)) || (rc = dalloc(c, (void**)&a.dinv, fb)) ||
                (rc = dalloc(c, &a.r, fb)) || (rc = dalloc(c, &a.z, fb)) || (rc = dalloc(c, &a.t, fb)) ||
                (l == 1 && (rc = dalloc(c, &a.w, fb))))
                return rc;
        }
        c->mg_alloc = true;
    }
    if (f16 && !c->mg16_alloc) {
        const size_t hb = (size_t)c->g.n * 2;
        if ((rc = dalloc(c, &c->bx16, hb)) || (rc = dalloc(c, &c->by16, hb)) || (rc = dalloc(c, &c->bz16, hb)) ||
            (rc = dalloc(c, &c->dinv16, hb)))
            return rc;
        for (size_t l = 1; l < c->levels.size(); l++) {
            CoarseArrays& a = c->coarse[l - 1];
            const size_t n2 = (size_t)c->levels[l].n * 2;
            if ((rc = dalloc(c, &a.bx16, n2)) || (rc = dalloc(c, &a.by16, n2)) || (rc = dalloc(c, &a.bz16, n2)) ||
                (rc = dalloc(c, &a.react16, n2)) || (rc = dalloc(c, &a.dinv16, n2)))
                return rc;
        }
        c->mg16_alloc = true;
    }
    return MHDP_OK;
}

// Frees every multigrid array (coarse levels and the BF16 copies of the fine coefficients).
void release_mg(mhdp_ctx* c) {
    (void)hipStreamSynchronize(c->stream);
    for (CoarseArrays& a : c->coarse) {
        void* q[] = {a.bx, a.by, a.bz, a.react, a.diag, a.dinv, a.bx16, a.by16, a.bz16, a.react16, a.dinv16,
                     a.r, a.z, a.t, a.w};
        for (void* p : q)
            if (p) (void)hipFree(p);
    }
    c->coarse.clear();
    void** fine16[] = {&c->bx16, &c->by16, &c->bz16, &c->dinv16};
    for (void** p : fine16) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    c->levels.clear();
    c->bytes -= c->mg_bytes;
    c->mg_bytes = 0;
    c->mg_alloc = c->mg16_alloc = c->fine16_fresh = false;
}

// The optional multigrid storage (Multigrid::new, src/fluid/multigrid.rs:85-96; attempted once per
// context, src/fluid/mod.rs:1035-1043): true when it is (or now is) allocated; a failed attempt
// frees what it had allocated and marks the context, whose solves then use Jacobi.
bool ensure_mg(mhdp_ctx* c, bool f16) {
    if (c->mg_refused) return false;
    if (c->mg_alloc && (!f16 || c->mg16_alloc)) return true;
    if (alloc_mg(c, f16) == MHDP_OK) return true;
    (void)hipGetLastError();
    release_mg(c);
    c->mg_refused = true;
    c->err.clear();
    return false;
}

// Galerkin coarse graphs of all levels (Multigrid::build, multigrid.rs:266-273) and, for
// the mixed solver, their BF16 copies.
void build_mg(mhdp_ctx* c, int meanfree, bool f16) {
    for (size_t l = 1; l < c->levels.size(); l++) {
        CoarseArrays& a = c->coarse[l - 1];
        const bool fine = l == 1;
        const float* fbx = fine ? c->bx : c->coarse[l - 2].bx;
        const float* fby = fine ? c->by : c->coarse[l - 2].by;
        const float* fbz = fine ? c->bz : c->coarse[l - 2].bz;
        const float* fdinv = fine ? c->dinv : c->coarse[l - 2].dinv;
        const float* freact = fine ? nullptr : c->coarse[l - 2].react;
        void* const c16[5] = {a.bx16, a.by16, a.bz16, a.react16, a.dinv16};  // the graph pass writes them too
        launch_mg_build(c->stream, c->levels[l], c->levels[l - 1], fine, meanfree, c->mask, fdinv, fbx, fby, fbz,
                        freact, a.bx, a.by, a.bz, a.react, a.diag, a.dinv, f16 ? c16 : nullptr);
    }
    if (f16 && !c->fine16_fresh) {  // the coefficient pass did not write the BF16 copies
        launch_to_bf16(c->stream, c->g.n, c->bx, c->bx16);
        launch_to_bf16(c->stream, c->g.n, c->by, c->by16);
        launch_to_bf16(c->stream, c->g.n, c->bz, c->bz16);
        launch_to_bf16(c->stream, c->g.n, c->dinv, c->dinv16);
        c->fine16_fresh = true;
    }
}

std::vector<LevelDev> level_view(mhdp_ctx* c) {
    std::vector<LevelDev> lv(c->levels.size());
    LevelDev& f = lv[0];
    f.g = c->g;
    f.bx = c->bx; f.by = c->by; f.bz = c->bz; f.react = nullptr; f.diag = c->diag; f.dinv = c->dinv;
    f.bx16 = c->bx16; f.by16 = c->by16; f.bz16 = c->bz16; f.react16 = nullptr; f.dinv16 = c->dinv16;
    f.r = c->r; f.z = c->z; f.t = c->t;
    for (size_t l = 1; l < c->levels.size(); l++) {
        CoarseArrays& a = c->coarse[l - 1];
        LevelDev& d = lv[l];
        d.g = c->levels[l];
        d.bx = a.bx; d.by = a.by; d.bz = a.bz; d.react = a.react; d.diag = a.diag; d.dinv = a.dinv;
        d.bx16 = a.bx16; d.by16 = a.by16; d.bz16 = a.bz16; d.react16 = a.react16; d.dinv16 = a.dinv16;
        d.r = a.r; d.z = a.z; d.t = a.t; d.w = a.w;
    }
    return lv;
}

int read_scalars(mhdp_ctx* c) {
    HIP_OK(hipMemcpyAsync(c->sc_h, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

// convergence(), mod.rs:969-977
int convergence(mhdp_ctx* c, float* residual, bool* conv, uint32_t* iters) {
    int rc = read_scalars(c);
    if (rc) return rc;
    const Scal& s = *c->sc_h;
    if (s.fault) return kBreakdown;
    const double res = std::sqrt(s.rr / std::max(s.bb, 1e-38));
    const double v[8] = {s.rz, s.rr, s.bb, s.sr, s.alpha, s.beta, s.conv, s.iters};
    bool bad = !std::isfinite(res) || s.rr < 0.0 || s.bb < 0.0;
    for (double x : v) bad |= !std::isfinite(x);
    if (bad) return fail(c, MHDP_NONFINITE, "Nonfinite fluid pressure residual");
    *residual = (float)res;
    *conv = s.conv > 0.5;
    *iters = (uint32_t)s.iters;
    return MHDP_OK;
}

// MHDP_STORED_MOBILITY=1: the operator apply and the FP64 residual pass load the stored mobilities
// (comparison runs).
bool formed_mobility_env() {
    static const bool on = [] {
        const char* e = std::getenv("MHDP_STORED_MOBILITY");
        return !(e && e[0] == '1');
    }();
    return on;
}

struct Solve {
    bool meanfree = false;
    bool mg = false;
    bool fused = false;  // fused fine-grid passes (multigrid with at least one coarse level)
    int store = STORE_F32;
    bool coarse64 = false;
    bool stall = false;  // a stalled inner solve counts as a breakdown (tuned cycles)
    std::vector<LevelDev> lv;
};

void refresh_stages(mhdp_ctx* c, const Solve& S) {  // mod.rs:1004-1014
    launch_refresh(c->stream, c->g, c->x, c->b, c->bx, c->by, c->bz, c->dinv, c->r, c->p, c->part, c->sc,
                   S.meanfree, 0, S.meanfree ? FIN_REFRESH_MEAN : FIN_INITIAL);
    if (S.meanfree) launch_subspace(c->stream, c->g, c->r, c->p, c->dinv, c->part, c->sc, 1, FIN_INITIAL_SUB);
}

// V-cycle with the fine level fused into two passes; its last pass reduces r.z (`mode`).
// `update` (brick path): the first pass also applies the PCG residual update.
void vcycle_fused(mhdp_ctx* c, Solve& S, int mode, bool update = false) {
    hipStream_t s = c->stream;
    const bool f16 = S.store == STORE_BF16;
    const void* dinv = f16 ? c->dinv16 : (const void*)c->dinv;
    const void* bx = f16 ? c->bx16 : (const void*)c->bx;
    const void* by = f16 ? c->by16 : (const void*)c->by;
    const void* bz = f16 ? c->bz16 : (const void*)c->bz;
    LevelDev* lv = S.lv.data();
    const int L = (int)S.lv.size();
    const bool brick = c->brick_enabled && brick_ok(c->g);
    const bool rb = c->rb && rb_level_ok(c->g);  // red-black passes on the fine level
    c->timer.before(MHDP_K_RESTRICT, s);
    const bool march = brick && c->march_enabled && S.store == STORE_BF16 && march_ok(c->g);
    if (rb)
        launch_rb_restrict(s, S.store, false, lv[1].g, lv[0].g, c->r, c->r2, c->r, c->r2, update ? c->q : nullptr, dinv,
                           bx, by, bz, lv[1].r, c->part, c->sc);
    else if (march && update)
        launch_march_restrict_update(s, lv[1].g, lv[0].g, c->r, c->r2, c->q, dinv, bx, by, bz, lv[1].r,
                                     c->part, c->sc);
    else if (brick && update)
        launch_brick_restrict_update(s, S.store, lv[1].g, lv[0].g, c->r, c->r2, c->q, dinv, bx, by, bz, lv[1].r,
                                     c->part, c->sc);
    else if (brick)
        launch_brick_restrict(s, S.store, lv[1].g, lv[0].g, c->r, dinv, bx, by, bz, lv[1].r, c->sc);
    else
        launch_fused_restrict(s, S.store, lv[1].g, lv[0].g, c->r, dinv, bx, by, bz, lv[1].r, c->sc);
    c->timer.after(MHDP_K_RESTRICT, s);
    void* z1 = nullptr;
    launch_vcycle_coarse(s, S.store, L, lv, c->sc, S.coarse64, c->coarse_visits, &z1, c->rb);
    record_mark();
    c->timer.before(MHDP_K_SMOOTH, s);
    if (rb)
        launch_rb_smooth(s, S.store, false, lv[0].g, lv[1].g, c->r, c->r2, dinv, bx, by, bz, z1, lv[0].t, nullptr, 0.0f,
                         c->part, c->sc, mode, update);
    else if (march)
        launch_march_smooth(s, lv[0].g, lv[1].g, c->r, c->r2, dinv, bx, by, bz, z1, lv[0].t, c->part, c->sc,
                            mode, update);
    else if (brick)
        launch_brick_smooth(s, S.store, lv[0].g, lv[1].g, c->r, c->r2, dinv, bx, by, bz, z1, lv[0].t, c->part, c->sc,
                            mode, update);
    else
        launch_fused_smooth(s, S.store, lv[0].g, lv[1].g, c->r, dinv, bx, by, bz, z1, lv[0].t, c->part, c->sc, mode);
    c->timer.after(MHDP_K_SMOOTH, s);
}

// The x update the fused path still owes (before any use of x).
void flush_x(mhdp_ctx* c, const Solve& S) {
    if (!S.fused) return;
    launch_fused_xflush(c->stream, c->g, c->p, c->p2, c->x, c->dinv, c->sc, S.meanfree ? 1 : 0);
    launch_clear_pending(c->stream, c->sc);
}

void mg_initialize(mhdp_ctx* c, Solve& S) {  // multigrid.rs:287-291
    if (S.fused) {
        vcycle_fused(c, S, FIN_INITIAL_MG);
        return;
    }
    void* zf = nullptr;
    launch_vcycle(c->stream, S.store, (int)S.lv.size(), S.lv.data(), c->r, c->mask, c->sc, S.coarse64, c->coarse_visits,
                  &zf);
    launch_pdot(c->stream, c->g, S.store, c->r, zf, c->dinv, c->p, c->part, c->sc, 1, FIN_INITIAL_MG);
}

void iterate_once(mhdp_ctx* c, Solve& S) {  // mod.rs:1057-1075, multigrid.rs:302-326
    hipStream_t s = c->stream;
    record_mark();
    if (S.fused) {
        c->timer.before(MHDP_K_APPLY, s);
        const bool brick = c->brick_enabled && brick_ok(c->g);
        // the mobilities formed in the pass from the liquid fraction and the open bytes (the
        // coefficient pass's values bit for bit) where both are at hand
        const bool formed = brick && c->open8_fresh && (!c->phase.two_phase || c->has_alpha) && formed_mobility_env();
        if (formed)
            launch_brick_apply_m(s, c->g, S.store, S.lv[0].t, c->p, c->p2, c->x,
                                 c->phase.two_phase ? c->alpha : nullptr, c->open8, c->phase.rho_liquid,
                                 c->phase.rho_air, c->q, c->part, c->sc, S.meanfree ? 1 : 0);
        else if (brick)
            launch_brick_apply(s, c->g, S.store, S.lv[0].t, c->p, c->p2, c->x, c->bx, c->by, c->bz,
                               S.store == STORE_BF16 ? c->dinv16 : (const void*)c->dinv, c->q, c->part, c->sc,
                               S.meanfree ? 1 : 0);
        else
            launch_fused_apply(s, c->g, S.store, S.lv[0].t, c->p, c->p2, c->x, c->bx, c->by, c->bz, c->dinv, c->q,
                               c->part, c->sc, S.meanfree ? 1 : 0);
        c->timer.after(MHDP_K_APPLY, s);
        if (!brick) {
            c->timer.before(MHDP_K_UPDATE, s);
            launch_fused_update(s, c->g, c->r, c->q, c->dinv, c->part, c->sc);
            c->timer.after(MHDP_K_UPDATE, s);
        }
        vcycle_fused(c, S, FIN_BETA_MG, brick);
        return;
    }
    c->timer.before(MHDP_K_APPLY, s);
    launch_apply(s, c->g, c->p, c->bx, c->by, c->bz, c->q, c->part, c->sc);
    c->timer.after(MHDP_K_APPLY, s);
    const int upd = S.meanfree ? FIN_MEAN : (S.mg ? FIN_STASH_RR : FIN_BETA);
    c->timer.before(MHDP_K_UPDATE, s);
    launch_update(s, c->g, c->x, c->r, c->p, c->q, c->dinv, c->part, c->sc, S.meanfree, upd);
    c->timer.after(MHDP_K_UPDATE, s);
    if (S.meanfree) launch_subspace(s, c->g, c->r, c->p, c->dinv, c->part, c->sc, 0, S.mg ? FIN_STASH_RR : FIN_BETA);
    if (S.mg) {
        void* zf = nullptr;
        launch_vcycle(s, S.store, (int)S.lv.size(), S.lv.data(), c->r, c->mask, c->sc, S.coarse64, c->coarse_visits, &zf,
                      &c->timer);
        launch_pdot(s, c->g, S.store, c->r, zf, c->dinv, c->p, c->part, c->sc, 0, FIN_BETA_MG);
        launch_psearch(s, c->g, S.store, zf, c->dinv, c->p, c->sc);
    } else {
        launch_search(s, c->g, c->r, c->dinv, c->p, c->sc);
    }
}

constexpr double kTunedOmega = 0.9;   // smoother weight picked by the V-cycle tuner
constexpr double kTunedScale = 1.6;   // coarse-correction scale picked by the V-cycle tuner (Jacobi)
constexpr double kTunedScaleRB = 1.8; // the same with red-black Gauss-Seidel smoothing
// ... and on grids of at least 2^24 cells, whose deep hierarchies keep the red-black cycle's
// level-1 spectrum well inside the positive-definite range (512^3: 12 instead of 13 iterations)
constexpr double kTunedScaleRBLarge = 1.9;
// The automatic smoother: red-black Gauss-Seidel on grids of at least 2^15 cells (64^3: 14 instead
// of 21 iterations, 1.60 instead of 1.81 ms; 128^3: 2.7 instead of 3.3 ms; 512^3: 12 instead of
// 21 iterations; the 64-problem batch of 32^3..96^3 grids: 20.1 instead of 20.6 ms), weighted
// Jacobi on smaller grids, whose few levels are all at the launch floor.
constexpr uint32_t kRedBlackAutoCells = 1u << 15;
// Grids of at least 2^24 cells take the larger red-black coarse scale.
constexpr uint32_t kRedBlackLargeCells = 1u << 24;
// Level-1 visits per cycle when the caller leaves the choice (coarse_visits = 0): a second visit
// (W-cycle at level 1) saves about a third of the iterations, but its coarse levels are launch-bound
// small kernels; measured on the plate workload it pays from 256^3 (12.1 vs 12.7 ms; 512^3: 75 vs 83
// ms) and costs 7-20% below 192^3.
constexpr uint32_t kTwoVisitMinCells = 1u << 24;
int coarse_visits_for(const mhdp_ctx* c, int requested) {
    if (requested > 0) return requested;
    return c->g.n >= kTwoVisitMinCells ? 2 : 1;
}

void set_solve_params(mhdp_ctx* c, Scal& s, bool meanfree, double tol2, double abs2, double bound,
                      double omega = kTunedOmega, double pscale = kTunedScale) {
    std::memset(&s, 0, sizeof s);
    s.omega = omega;
    s.pscale = pscale;
    s.meanfree = meanfree ? 1.0 : 0.0;
    s.tol2 = tol2;
    s.abs2 = abs2;
    s.bound = bound;
    s.inv_nf = meanfree ? 1.0 / (double)std::max<uint32_t>(c->fluid_cells, 1) : 0.0;
    s.anchor = c->anchor0;
}

// Iterations to issue before the next host read: as many as the residual reduction measured since
// (rr_prev, it_prev) predicts are left to reach the stopping bound, at least 1 and at most 8.  A host
// read costs about as much as a few launches that would only find the solve converged.
uint32_t next_batch(double rr_prev, uint32_t it_prev, const Scal& v) {
    const double target = (v.tol2 * v.bb + v.abs2) / std::max(v.bound, 1e-300);
    const uint32_t it = (uint32_t)v.iters;
    if (it <= it_prev || !(rr_prev > 0.0) || !(v.rr > 0.0) || !(target > 0.0) || v.rr >= rr_prev) return 8u;
    const double per = std::log(v.rr / rr_prev) / (double)(it - it_prev);  // log reduction per iteration
    const double left = std::log(target / v.rr) / per;
    if (!(left > 1.0)) return 1u;
    return (uint32_t)std::min(8.0, std::ceil(left));
}

// The fused iteration applies the operator with the brick kernels, whose first iteration of a
// solve reads no old search direction: the starts then leave p and p2 unset.
bool brick_apply(const mhdp_ctx* c) { return c->brick_enabled && brick_ok(c->g); }
float* start_p(mhdp_ctx* c, float* p) { return brick_apply(c) ? nullptr : p; }

// The PCG loop of FluidState::project (mod.rs:1015-1104) on the FP32 vectors x (from 0),
// right-hand side c->b.  `reference_checks`: refresh on convergence and the independent
// physical check every 64 iterations, as the reference; the inner solves of the mixed
// solver skip them (the FP64 outer residual is the acceptance test).
// `want_mg`: build the multigrid preconditioner lazily, only once a nonzero residual
// needs it (mod.rs:1031-1049).
int pcg(mhdp_ctx* c, Solve& S, bool want_mg, uint32_t max_iter, bool reference_checks, uint32_t* iterations,
        float* residual) {
    launch_fill_f32(c->stream, c->g.n, c->x, 0.0f);
    refresh_stages(c, S);
    bool conv = false;
    uint32_t seen = 0, it = 0;
    int rc = convergence(c, residual, &conv, &seen);
    if (rc) return rc;
    if (!conv && want_mg && !S.mg && ensure_mg(c, S.store == STORE_BF16)) {
        build_mg(c, S.meanfree ? 1 : 0, S.store == STORE_BF16);
        S.mg = true;
        S.lv = level_view(c);
        S.fused = S.lv.size() >= 2 && c->fused_enabled;
    }
    if (!conv && S.fused && !brick_apply(c)) {  // finite search directions for the first p = z + 0 p
        launch_fill_f32(c->stream, c->g.n, c->p, 0.0f);
        launch_fill_f32(c->stream, c->g.n, c->p2, 0.0f);
    }
    if (!conv && S.mg) mg_initialize(c, S);
    // multigrid: batches of 8, then as many as the measured convergence rate predicts are left
    uint32_t want = S.mg ? 8u : 32u, it_prev = 0;
    double rr_prev = c->sc_h->rr;
    while (!conv && it < max_iter) {
        uint32_t batch = want;
        batch = std::min(batch, max_iter - it);
        batch = std::min(batch, 64u - it % 64u);
        for (uint32_t k = 0; k < batch; k++) iterate_once(c, S);
        if ((rc = convergence(c, residual, &conv, &seen))) return rc;
        if (seen <= it && !conv) return fail(c, MHDP_NO_PROGRESS, "Fluid pressure iteration made no progress");
        it = seen;
        if (S.mg && !conv) {
            want = next_batch(rr_prev, it_prev, *c->sc_h);
            rr_prev = c->sc_h->rr;
            it_prev = it;
        }
        if (conv && reference_checks) {
            flush_x(c, S);
            refresh_stages(c, S);
            if ((rc = convergence(c, residual, &conv, &seen))) return rc;
            if (!conv && S.mg) mg_initialize(c, S);
        } else if (!conv && reference_checks && (it % 64u == 0u || it == max_iter)) {
            flush_x(c, S);
            launch_refresh(c->stream, c->g, c->x, c->b, c->bx, c->by, c->bz, c->dinv, c->r, c->p, c->part, c->sc,
                           S.meanfree, 1, FIN_CHECK);
            if ((rc = convergence(c, residual, &conv, &seen))) return rc;
        }
    }
    flush_x(c, S);
    *iterations = it;
    return conv ? MHDP_OK : MHDP_NOT_CONVERGED;
}


int upload_scal(mhdp_ctx* c, const Scal& s) {
    *c->sc_h = s;
    HIP_OK(hipMemcpyAsync(c->sc, c->sc_h, sizeof(Scal), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

// The per-cell open-face bytes when they match the mask (null otherwise).
const uint8_t* open8_of(const mhdp_ctx* c) { return c->open8_fresh ? c->open8 : nullptr; }

// The FP64 residual pass that starts a refinement step (its stored-mobility form where the
// mobilities cannot be formed in the pass).
void outer_start(mhdp_ctx* c, hipStream_t s, bool meanfree, bool zero_x) {
    const bool formable = (!c->phase.two_phase || c->has_alpha) && formed_mobility_env();
    if (formable && launch_outer_start_m(s, c->g, c->x64, c->b64, c->phase.two_phase ? c->alpha : nullptr,
                                         open8_of(c), c->phase.rho_liquid, c->phase.rho_air, meanfree, c->r, c->x,
                                         start_p(c, c->p), start_p(c, c->p2), c->part, c->sc, zero_x))
        return;
    launch_outer_start(s, c->g, c->x64, c->b64, c->bx, c->by, c->bz, c->dinv, c->r, c->x, start_p(c, c->p),
                       start_p(c, c->p2), c->part, c->sc, zero_x);
}

// FP64 iterative refinement over the fused multigrid path.  Each outer step is one residual
// pass that also starts the inner solve on the device (inner tolerance, deferred mean-free
// projection, zeroed x and search directions), the inner V-cycle PCG in batches (8, then as
// many as the measured convergence rate predicts are left), and one pass folding the inner
// solution (with its owed last update) into x64.  The host reads the scalars after each batch,
// and after each later step's residual pass: a step whose FP64 residual already meets the
// target launches nothing more.
int refine_fused(mhdp_ctx* c, Solve& S, const mhdp_opts* o, uint32_t max_outer, bool warm, bool started,
                 uint32_t* iterations, uint32_t* outer, double* res64) {
    hipStream_t s = c->stream;
    *iterations = 0;
    *outer = 0;
    auto finite = [](const Scal& v) {
        const double vals[10] = {v.rr64, v.bb64, v.rz, v.rr, v.bb, v.sr, v.alpha, v.beta, v.conv, v.iters};
        for (double x : vals)
            if (!std::isfinite(x)) return false;
        return true;
    };
    for (bool first = true;; first = false) {
        // `started`: the caller has run the first step's residual pass and read its scalars
        if (!(first && started)) {
            outer_start(c, s, S.meanfree, first && !warm);
            int rc = read_scalars(c);
            if (rc) return rc;
        }
        {
            const Scal& v = *c->sc_h;
            if (!finite(v)) return fail(c, MHDP_NONFINITE, "Nonfinite fluid pressure residual");
            *res64 = v.bb64 > 0.0 ? std::sqrt(v.rr64 / v.bb64) : 0.0;
            if (v.outer_done) {
                if (first && !warm) launch_fill_f64(s, c->g.n, c->x64, 0.0);
                return MHDP_OK;
            }
            if (*outer >= max_outer) return MHDP_NOT_CONVERGED;
        }
        mg_initialize(c, S);
        uint32_t it = 0, it_prev = 0, batch = 8u, best_it = 0;
        double rr_prev = c->sc_h->rr;  // the inner solve's initial |r|^2
        double best = rr_prev;
        for (;;) {
            batch = std::min(batch, o->max_iter - it);
            for (uint32_t k = 0; k < batch; k++) iterate_once(c, S);
            int rc = read_scalars(c);
            if (rc) return rc;
            const Scal& v = *c->sc_h;
            if (v.fault) return kBreakdown;
            if (!finite(v)) return fail(c, MHDP_NONFINITE, "Nonfinite fluid pressure residual");
            const bool conv = v.conv > 0.5;
            const uint32_t seen = (uint32_t)v.iters;
            if (seen <= it && !conv) return fail(c, MHDP_NO_PROGRESS, "Fluid pressure iteration made no progress");
            it = seen;
            if (conv || it >= o->max_iter) break;
            if (v.rr < best) {
                best = v.rr;
                best_it = it;
            } else if (S.stall && it - best_it >= kStallIterations) {
                return kBreakdown;
            }
            batch = next_batch(rr_prev, it_prev, v);
            rr_prev = v.rr;
            it_prev = it;
        }
        *iterations += it;
        launch_outer_update(s, c->g, c->x64, c->x, c->p, c->p2, c->dinv, c->sc, S.meanfree, first && !warm);
        (*outer)++;
    }
}

int ensure_fine(mhdp_ctx* c) {
    const size_t n = c->g.n, fb = n * sizeof(float);
    int rc;
    if ((rc = dalloc(c, (void**)&c->mask, n * 4)) || (rc = dalloc(c, (void**)&c->alpha, fb)) ||
        (rc = dalloc(c, (void**)&c->bx, fb)) || (rc = dalloc(c, (void**)&c->by, fb)) ||
        (rc = dalloc(c, (void**)&c->bz, fb)) || (rc = dalloc(c, (void**)&c->diag, fb)) ||
        (rc = dalloc(c, (void**)&c->dinv, fb)) || (rc = dalloc(c, (void**)&c->b, fb)) ||
        (rc = dalloc(c, (void**)&c->x, fb)) || (rc = dalloc(c, (void**)&c->r, fb)) ||
        (rc = dalloc(c, (void**)&c->p, fb)) || (rc = dalloc(c, (void**)&c->q, fb)) ||
        (rc = dalloc(c, (void**)&c->p2, fb)) || (rc = dalloc(c, (void**)&c->r2, fb)) ||
        (rc = dalloc(c, &c->z, fb)) || (rc = dalloc(c, &c->t, fb)) ||
        (rc = dalloc(c, (void**)&c->x64, n * 8)) || (rc = dalloc(c, (void**)&c->b64, n * 8)) ||
        (rc = dalloc(c, (void**)&c->phi_prev, n * 8)) ||
        (rc = dalloc(c, (void**)&c->sc, sizeof(Scal))) ||
        (rc = dalloc(c, (void**)&c->part, kMaxPartials * 4 * sizeof(double))))
        return rc;
    return MHDP_OK;
}

PhaseP phase_params(const mhdp_ctx* c, bool meanfree) {
    PhaseP ph;
    ph.two_phase = c->phase.two_phase ? 1 : 0;
    ph.rho_l = c->phase.rho_liquid;
    ph.rho_air = c->phase.rho_air;
    ph.meanfree = meanfree ? 1 : 0;
    return ph;
}

// Face mobilities, diagonal and 1/diag of this projection; `want16`: also their BF16 copies
// for the mixed solver's V-cycle (when allocated).
int prepare(mhdp_ctx* c, bool meanfree, bool want16 = false) {
    if (!c->cells) return fail(c, MHDP_INVALID, "Fluid cells have not been set");
    const bool h = want16 && c->mg16_alloc;
    c->fine16_fresh = launch_coefficients(c->stream, c->g, phase_params(c, meanfree), c->mask, open8_of(c),
                                          c->has_alpha ? c->alpha : nullptr, c->bx, c->by, c->bz, c->diag, c->dinv,
                                          h ? c->bx16 : nullptr, h ? c->by16 : nullptr, h ? c->bz16 : nullptr,
                                          h ? c->dinv16 : nullptr);
    return MHDP_OK;
}

// multigrid.wgsl:69,87 (the weight is the f32 2/3 of the WGSL); multigrid.rs:67-82
constexpr Cycle kReferenceCycle{(double)(2.0f / 3.0f), 1.0, 1, false};
bool is_reference(const Cycle& cy) {
    return !cy.rb && cy.omega == kReferenceCycle.omega && cy.pscale == kReferenceCycle.pscale && cy.visits == 1;
}

// The cycle the options ask for: smoother, its weight and coarse scale (0 = tuned defaults), visits.
Cycle requested_cycle(const mhdp_ctx* c, const mhdp_opts* o) {
    Cycle cy;
    cy.rb = o->smoother == MHDP_SMOOTH_RED_BLACK ||
            (o->smoother == MHDP_SMOOTH_AUTO && c->g.n >= kRedBlackAutoCells);
    cy.omega = o->smooth_weight > 0.0f ? (double)o->smooth_weight : kTunedOmega;
    cy.pscale = o->coarse_scale > 0.0f ? (double)o->coarse_scale
                                       : (cy.rb ? (c->g.n >= kRedBlackLargeCells ? kTunedScaleRBLarge : kTunedScaleRB)
                                                : kTunedScale);
    cy.visits = coarse_visits_for(c, o->coarse_visits);
    return cy;
}

// One projection with the cycle `cy`; kBreakdown when its multigrid PCG broke down.
int project_once(mhdp_ctx* c, const float* ustar, const mhdp_opts* o, const Cycle& cy, float* uout, double* phi,
                 mhdp_stats* st) {
    int rc;
    const bool meanfree = o->gauge != MHDP_GAUGE_PINNED && c->components == 1;
    const bool mixed = o->precision == MHDP_MIXED;
    const bool preferred = o->precond == MHDP_PRECOND_MG || (o->precond == MHDP_PRECOND_AUTO && c->g.n >= kMgMinCells);
    const bool warm = o->warm_start && c->prev_valid;
    if (warm && !mixed && (rc = dalloc(c, (void**)&c->uwarm, face_count(c->grid) * sizeof(float)))) return rc;

    const double omega = cy.omega, pscale = cy.pscale;
    c->coarse_visits = cy.visits;
    c->rb = cy.rb;
    Solve S;
    S.meanfree = meanfree;
    S.store = mixed ? STORE_BF16 : STORE_F32;
    S.coarse64 = mixed;
    S.stall = !is_reference(cy);
    const float nf = (float)std::max<uint32_t>(c->fluid_cells, 1);

    c->timer.used = 0;
    HIP_OK(hipEventRecord(c->ev0, c->stream));
    if ((rc = prepare(c, meanfree, mixed && preferred))) return rc;
    if (warm && !mixed) {  // pressure_warm: the solve then finds the correction to the previous potential
        launch_warm(c->stream, c->g, c->grid.h, c->mask, c->phi_prev, c->bx, c->by, c->bz, ustar, c->uwarm);
        ustar = c->uwarm;
    }
    uint32_t iterations = 0, outer = 0;
    float residual = 0.0f;
    double res64 = 0.0;
    bool phi_copied = false;  // the velocity correction also wrote the potential to `phi`
    if (!mixed) {
        // Params.solve (mod.rs:398-404, 978-1002) in FP32 as the reference computes it
        const float tol = o->rel_tol, ab = o->abs_div_tol, h = c->grid.h;
        const float h2 = h * h;
        const float tol2 = tol * tol;
        const float abs2 = ((ab * ab) * (h2 * h2)) * (float)c->fluid_cells;
        Scal s;
        set_solve_params(c, s, meanfree, tol2, abs2, meanfree ? 1.0 : (double)nf, omega, pscale);
        s.guard = preferred ? 1u : 0u;
        if ((rc = upload_scal(c, s))) return rc;
        launch_rhs_f32(c->stream, c->g, c->grid.h, c->mask, open8_of(c), c->bx, c->by, c->bz, ustar, c->b);
        rc = pcg(c, S, preferred, o->max_iter, true, &iterations, &residual);
        if (rc == MHDP_NOT_CONVERGED) {
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "Fluid pressure solve did not converge after %u iterations (true relative residual "
                          "%.3e, requested %.3e, RMS divergence floor %.3e s\xe2\x81\xbb\xc2\xb9)",
                          iterations, (double)residual, (double)o->rel_tol, (double)o->abs_div_tol);
            return fail(c, rc, buf);
        }
        if (rc) return rc;
        if (uout) launch_correct_f32(c->stream, c->g, c->grid.h, c->bx, c->by, c->bz, c->x, ustar, uout);
        launch_cast_f32_f64(c->stream, c->g.n, c->x, c->x64);
    } else {
        // FP64 iterative refinement: residual in FP64, inner FP32 PCG with the BF16 V-cycle.
        const double tau = o->inner_tol > 0.0f ? (double)o->inner_tol : 1e-4;
        const uint32_t max_outer = o->max_outer ? o->max_outer : 16u;
        Scal s;
        set_solve_params(c, s, meanfree, tau * tau, 0.0, 1.0, omega, pscale);
        s.refine_tol = o->refine_tol;
        s.tau = tau;
        s.guard = preferred ? 1u : 0u;
        if ((rc = upload_scal(c, s))) return rc;
        // The first refinement step's FP64 residual decides whether the optional multigrid is
        // needed at all (a zero or already-met residual never allocates it, mod.rs:1031-1043).
        // From a zero potential it is the right-hand side itself: one pass forms both.
        if (warm || !launch_rhs_start(c->stream, c->g, c->grid.h, open8_of(c), ustar, c->b64, meanfree, c->r, c->x,
                                      start_p(c, c->p), start_p(c, c->p2), c->part, c->sc)) {
            launch_rhs_f64(c->stream, c->g, c->grid.h, c->mask, open8_of(c), c->bx, c->by, c->bz, ustar, c->b64);
            // a warm start is the refinement begun from the previous potential (the FP64 residual
            // then sees exactly b - L phi_old; faces are corrected from u* with the full potential)
            if (warm) launch_old_potential(c->stream, c->g.n, c->mask, c->phi_prev, c->x64);
            outer_start(c, c->stream, meanfree, !warm);
        }
        if ((rc = read_scalars(c))) return rc;
        const double r0 = c->sc_h->rr64, b0 = c->sc_h->bb64;
        if (!std::isfinite(r0) || !std::isfinite(b0)) return fail(c, MHDP_NONFINITE, "Nonfinite fluid pressure residual");
        bool done = c->sc_h->outer_done != 0u;
        const bool fast = !done && preferred && c->fused_enabled && ensure_mg(c, true) && c->levels.size() >= 2;
        if (done) {
            res64 = b0 > 0.0 ? std::sqrt(r0 / b0) : 0.0;
            if (!warm) launch_fill_f64(c->stream, c->g.n, c->x64, 0.0);
        } else if (fast) {
            build_mg(c, meanfree ? 1 : 0, true);
            S.mg = true;
            S.lv = level_view(c);
            S.fused = true;
            rc = refine_fused(c, S, o, max_outer, warm, true, &iterations, &outer, &res64);
            if (rc && rc != MHDP_NOT_CONVERGED) return rc;
            done = rc == MHDP_OK;
        } else if (!warm) {
            launch_fill_f64(c->stream, c->g.n, c->x64, 0.0);
        }
        for (; !fast && !done;) {
            launch_residual64(c->stream, c->g, c->x64, c->b64, c->bx, c->by, c->bz, c->dinv, c->b, c->part, c->sc);
            if ((rc = read_scalars(c))) return rc;
            const double rr = c->sc_h->rr, bb = c->sc_h->spare;
            if (!std::isfinite(rr) || !std::isfinite(bb)) return fail(c, MHDP_NONFINITE, "Nonfinite fluid pressure residual");
            res64 = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
            if (bb == 0.0 || res64 <= o->refine_tol) {
                done = true;
                break;
            }
            if (outer >= max_outer) break;
            // inner tolerance: no tighter than the remaining factor asks for
            double t = std::max(tau, 0.5 * o->refine_tol / res64);
            if (t > 0.5) t = 0.5;
            Scal si;
            set_solve_params(c, si, meanfree, t * t, 0.0, 1.0, omega, pscale);
            si.guard = preferred ? 1u : 0u;
            if ((rc = upload_scal(c, si))) return rc;
            uint32_t it = 0;
            rc = pcg(c, S, preferred, o->max_iter, false, &it, &residual);
            iterations += it;
            if (rc && rc != MHDP_NOT_CONVERGED) return rc;
            launch_xupdate64(c->stream, c->g, c->x64, c->x, c->dinv);
            outer++;
        }
        if (!done) {
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "Fluid pressure solve did not converge after %u iterations (true relative residual %.3e, "
                          "requested %.3e)",
                          iterations, res64, o->refine_tol);
            return fail(c, MHDP_NOT_CONVERGED, buf);
        }
        residual = (float)res64;
        if (uout) phi_copied = launch_correct_f64(c->stream, c->g, c->grid.h, c->bx, c->by, c->bz, c->x64, ustar,
                                                  uout, phi);
    }
    // the potential (plus the previous one after an FP32 warm start) is returned and kept
    if (warm && !mixed) launch_commit(c->stream, c->g.n, c->mask, c->phi_prev, c->x64);
    if (phi && !phi_copied)
        HIP_OK(hipMemcpyAsync(phi, c->x64, (size_t)c->g.n * 8, hipMemcpyDeviceToDevice, c->stream));
    std::swap(c->x64, c->phi_prev);
    c->prev_valid = true;
    HIP_OK(hipEventRecord(c->ev1, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipGetLastError());
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    // Launches after convergence inside a batch return at once; count only launches that did
    // the work (longer than a tenth of the longest one).
    c->timer_ms = 0.0;
    c->timer_launches = 0;
    std::vector<float> durations;
    for (size_t k = 0; k + 1 < c->timer.used; k += 2) {
        float t = 0.0f;
        HIP_OK(hipEventElapsedTime(&t, c->timer.ev[k], c->timer.ev[k + 1]));
        durations.push_back(t);
    }
    float longest = 0.0f;
    for (float t : durations) longest = std::max(longest, t);
    for (float t : durations)
        if (t >= 0.1f * longest) {
            c->timer_ms += t;
            c->timer_launches++;
        }
    if (st) {
        st->iterations = iterations;
        st->outer_steps = outer;
        st->components = c->components;
        st->fluid_cells = c->fluid_cells;
        st->levels = S.mg ? (uint32_t)c->levels.size() : 0u;
        st->meanfree = meanfree ? 1 : 0;
        st->multigrid = S.mg ? 1 : 0;
        st->rel_residual = residual;
        st->rel_residual_f64 = mixed ? res64 : -1.0;
        st->seconds = ms * 1e-3;
        st->cycle_fallback = 0;
    }
    return MHDP_OK;
}

// FluidState::project: validate, keep the FP64 vectors, solve with the requested (tuned) cycle;
// if its multigrid PCG breaks down (r.z or p.Ap <= 0: the tuned cycle was not positive definite on
// this problem), solve again with the reference's V-cycle, which is.
int project_impl(mhdp_ctx* c, const float* ustar, const mhdp_opts* o, float* uout, double* phi, mhdp_stats* st) {
    int rc = validate_opts(c, o);
    if (rc) return rc;
    if (!c->cells) return fail(c, MHDP_INVALID, "Fluid cells have not been set");
    if (!ustar) return fail(c, MHDP_INVALID, "Missing candidate face velocities");
    Cycle cy = requested_cycle(c, o);
    if (c->test_cycle.visits > 0) {  // test hook: the first attempt uses an unvalidated cycle
        cy = c->test_cycle;
        c->test_cycle = Cycle{0.0, 0.0, 0};
    }
    rc = project_once(c, ustar, o, cy, uout, phi, st);
    if (rc == kBreakdown && !is_reference(cy)) {
        rc = project_once(c, ustar, o, kReferenceCycle, uout, phi, st);
        if (rc == MHDP_OK && st) st->cycle_fallback = 1;
    }
    if (rc == kBreakdown)
        return fail(c, MHDP_NO_PROGRESS, "Fluid pressure iteration made no progress (the preconditioner is not positive definite)");
    return rc;
}

// ============================================================================ batched projection
// Lockstep driver of many mixed-precision multigrid projections.  The launches of every stage of
// every problem are recorded (go<K> appends instead of launching), the per-problem sequences are
// merged kernel by kernel, and each merged group is ONE launch over all its problems (k_many).
// Block numbering, arithmetic and reduction order per problem are those of a single projection,
// so every result is bit-identical to mhdp_project_dev's.

struct MergedOp {
    void (*issue)(hipStream_t, const void*, const uint32_t*, uint32_t, uint32_t);
    uint32_t count, blocks;
    size_t args_off, starts_off;
};
struct Plan {
    std::vector<MergedOp> ops;
    void* dev = nullptr;
};
struct BatchMemory {
    std::vector<void*> blocks;
    ~BatchMemory() {
        for (void* p : blocks) (void)hipFree(p);
    }
    void* get(size_t bytes) {
        void* p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        blocks.push_back(p);
        return p;
    }
};

// Record f(p) for every member p, merge the sequences and upload the merged arguments.
template <class F>
bool record_plan(const std::vector<int>& members, BatchMemory& mem, Plan& plan, F&& f) {
    std::vector<Recorder> recs(members.size());
    for (size_t i = 0; i < members.size(); i++) {
        Recorder::active() = &recs[i];
        f(members[i]);
        Recorder::active() = nullptr;
    }
    plan.ops.clear();
    std::vector<unsigned char> blob;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<size_t> head(recs.size(), 0);
    for (;;) {
        // the kernel waiting at the most sequence heads goes next (ties: lowest problem)
        std::vector<std::pair<const void*, size_t>> kinds;
        bool any = false;
        for (size_t p = 0; p < recs.size(); p++) {
            if (head[p] >= recs[p].ops.size()) continue;
            any = true;
            const void* k = recs[p].ops[head[p]].kind;
            if (!k) continue;  // waiting at a mark
            auto it = std::find_if(kinds.begin(), kinds.end(), [&](const auto& e) { return e.first == k; });
            if (it == kinds.end()) kinds.push_back({k, 1});
            else it->second++;
        }
        if (!any) break;
        if (kinds.empty()) {  // every unfinished sequence is at a mark: pass it together
            for (size_t p = 0; p < recs.size(); p++)
                if (head[p] < recs[p].ops.size()) head[p]++;
            continue;
        }
        const void* best = kinds[0].first;
        size_t bestn = kinds[0].second;
        for (const auto& e : kinds)
            if (e.second > bestn) {
                best = e.first;
                bestn = e.second;
            }
        MergedOp m{};
        std::vector<uint32_t> starts{0u}, counts;
        m.args_off = align(blob.size());
        for (size_t p = 0; p < recs.size(); p++) {
            if (head[p] >= recs[p].ops.size() || recs[p].ops[head[p]].kind != best) continue;
            const Recorded& op = recs[p].ops[head[p]];
            if (blob.size() < m.args_off) blob.resize(m.args_off);
            blob.insert(blob.end(), op.args.begin(), op.args.end());
            // 8-aligned starts: each problem keeps its XCD pattern (see run_many)
            starts.push_back(starts.back() + (op.blocks + MHDP_XCDS - 1) / MHDP_XCDS * MHDP_XCDS);
            counts.push_back(op.blocks);
            m.issue = op.issue;
            head[p]++;
        }
        m.count = (uint32_t)counts.size();
        m.blocks = starts.back();
        m.starts_off = align(blob.size());
        blob.resize(m.starts_off);
        starts.insert(starts.end(), counts.begin(), counts.end());
        const unsigned char* sb = reinterpret_cast<const unsigned char*>(starts.data());
        blob.insert(blob.end(), sb, sb + starts.size() * sizeof(uint32_t));
        if (m.blocks) plan.ops.push_back(m);
    }
    plan.dev = mem.get(blob.size());
    if (!plan.dev) return false;
    return hipMemcpy(plan.dev, blob.data(), blob.size(), hipMemcpyHostToDevice) == hipSuccess;
}
void issue_plan(hipStream_t s, const Plan& plan) {
    const unsigned char* base = static_cast<const unsigned char*>(plan.dev);
    for (const MergedOp& m : plan.ops)
        m.issue(s, base + m.args_off, reinterpret_cast<const uint32_t*>(base + m.starts_off), m.count, m.blocks);
}

// A plan issued again and again with the same arguments (the rounds' 8 iterations), captured
// once into a HIP graph on a private stream and then submitted as one graph launch per round.
// The plan holds kernel launches only (no memset or copy nodes).  MHDP_BATCH_GRAPH=0 issues the
// launches one by one instead (comparison runs).
struct PlanGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~PlanGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    bool build(const Plan& plan) {
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return false;
        bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            issue_plan(cs, plan);
            ok = hipStreamEndCapture(cs, &graph) == hipSuccess && graph &&
                 hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        }
        (void)hipStreamDestroy(cs);
        if (!ok) (void)hipGetLastError();
        return ok;
    }
};
bool batch_graph_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("MHDP_BATCH_GRAPH");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Copies the scalar blocks of `count` problems into one array (read back with one copy).
struct KGatherScal {
    static constexpr int MINB = 0;
    struct Args {
        const Scal* const* src;
        uint32_t* dst;
        uint32_t count;
    };
    static __device__ void run(const Args& A, const VGrid& vg) {
        constexpr uint32_t words = sizeof(Scal) / 4;
        const uint32_t total = words * A.count;
        for (uint32_t i = vg.b * MHDP_BLOCK + threadIdx.x; i < total; i += vg.n * MHDP_BLOCK) {
            const uint32_t p = i / words, w = i - p * words;
            A.dst[i] = reinterpret_cast<const uint32_t*>(A.src[p])[w];
        }
    }
};

// Eligible for the lockstep driver: the fused multigrid path of the mixed solver.
// The batch allocates the multigrid storage before the first residual is known (it has no
// counterpart in the reference's lazily allocating FluidState); a context whose allocation is refused
// sends the whole batch to the per-context path, where that context solves with Jacobi.
bool lockstep_ok(mhdp_ctx* c, const mhdp_opts* o) {
    const bool preferred = o->precond == MHDP_PRECOND_MG || (o->precond == MHDP_PRECOND_AUTO && c->g.n >= kMgMinCells);
    return c->cells && o->precision == MHDP_MIXED && preferred && !o->warm_start && c->fused_enabled &&
           hierarchy(c->g).size() >= 2 && ensure_mg(c, true);
}

int project_lockstep(mhdp_ctx* const* ctxs, int count, const float* const* ustar, const mhdp_opts* o,
                     float* const* uout, double* const* phi, mhdp_stats* st, std::vector<int>& rcs) {
    const int n = count;
    mhdp_ctx* c0 = ctxs[0];
    const hipStream_t s = c0->stream;
    for (int p = 0; p < n; p++)
        if (hipStreamSynchronize(ctxs[p]->stream) != hipSuccess) return MHDP_DEVICE;
    struct PS {
        Solve S;
        bool running = false, done = false, inner = false;
        uint32_t outer = 0, iterations = 0, it = 0;
        uint32_t best_it = 0;  // iteration of the inner solve's least residual `best`
        double best = std::numeric_limits<double>::infinity();
        bool checked = false, breakdown = false;
        bool started = false;  // the set-up pass also ran the first step's residual pass
        double res64 = 0.0;
    };
    std::vector<PS> ps((size_t)n);
    // kernel timing is off during the batch; every exit restores each context's choice
    struct TimerRestore {
        mhdp_ctx* const* ctxs;
        std::vector<int> kernel;
        ~TimerRestore() {
            for (size_t p = 0; p < kernel.size(); p++) ctxs[p]->timer.kernel = kernel[p];
        }
    } timers{ctxs, std::vector<int>((size_t)n, -1)};
    const double tau = o->inner_tol > 0.0f ? (double)o->inner_tol : 1e-4;
    const uint32_t max_outer = o->max_outer ? o->max_outer : 16u;
    for (int p = 0; p < n; p++) {
        mhdp_ctx* c = ctxs[p];
        c->err.clear();
        int rc = validate_opts(c, o);
        Cycle cy = requested_cycle(c, o);
        if (c->test_cycle.visits > 0) {  // test hook (mhdp_probe_next_cycle)
            cy = c->test_cycle;
            c->test_cycle = Cycle{0.0, 0.0, 0};
        }
        if (!rc) {
            c->coarse_visits = cy.visits;
            c->rb = cy.rb;
        }
        if (!rc && !ustar[p]) rc = fail(c, MHDP_INVALID, "Missing candidate face velocities");
        PS& q = ps[(size_t)p];
        q.S.meanfree = o->gauge != MHDP_GAUGE_PINNED && c->components == 1;
        q.S.store = STORE_BF16;
        q.S.coarse64 = true;
        q.S.stall = !is_reference(cy);
        if (!rc) {
            Scal sc;
            set_solve_params(c, sc, q.S.meanfree, tau * tau, 0.0, 1.0, cy.omega, cy.pscale);
            sc.refine_tol = o->refine_tol;
            sc.tau = tau;
            sc.guard = 1u;
            rc = upload_scal(c, sc);
        }
        rcs[(size_t)p] = rc;
        q.running = rc == MHDP_OK;
        timers.kernel[(size_t)p] = c->timer.kernel;
        c->timer.kernel = -1;
        c->timer.used = 0;
    }
    auto members = [&](auto pred) {
        std::vector<int> m;
        for (int p = 0; p < n; p++)
            if (pred(ps[(size_t)p])) m.push_back(p);
        return m;
    };
    BatchMemory mem;
    hipEvent_t e0 = c0->ev0, e1 = c0->ev1;
    if (hipEventRecord(e0, s) != hipSuccess) return MHDP_DEVICE;
    // set-up: mobilities, diagonal, BF16 copies, FP64 RHS, Galerkin hierarchy
    {
        Plan plan;
        auto run = members([](const PS& q) { return q.running; });
        if (!record_plan(run, mem, plan, [&](int p) {
                mhdp_ctx* c = ctxs[p];
                PS& q = ps[(size_t)p];
                prepare(c, q.S.meanfree, true);
                // the first refinement step starts from a zero potential: its residual pass is the RHS pass
                q.started = launch_rhs_start(s, c->g, c->grid.h, open8_of(c), ustar[p], c->b64, q.S.meanfree, c->r,
                                             c->x, start_p(c, c->p), start_p(c, c->p2), c->part, c->sc);
                if (!q.started)
                    launch_rhs_f64(s, c->g, c->grid.h, c->mask, open8_of(c), c->bx, c->by, c->bz, ustar[p], c->b64);
                build_mg(c, q.S.meanfree ? 1 : 0, true);
                q.S.mg = true;
                q.S.lv = level_view(c);
                q.S.fused = true;
            }))
            return MHDP_NO_MEMORY;
        issue_plan(s, plan);
    }
    std::vector<const Scal*> srcs((size_t)n);
    for (int p = 0; p < n; p++) srcs[(size_t)p] = ctxs[p]->sc;
    void* src_d = mem.get(sizeof(Scal*) * (size_t)n);
    uint32_t* dst_d = static_cast<uint32_t*>(mem.get(sizeof(Scal) * (size_t)n));
    if (!src_d || !dst_d) return MHDP_NO_MEMORY;
    if (hipMemcpy(src_d, srcs.data(), sizeof(Scal*) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) return MHDP_DEVICE;
    std::vector<Scal> sh((size_t)n);
    auto read_all = [&]() -> bool {
        go<KGatherScal>(s, grid_blocks((uint32_t)(n * sizeof(Scal) / 4)), {(const Scal* const*)src_d, dst_d, (uint32_t)n});
        return hipMemcpyAsync(sh.data(), dst_d, sizeof(Scal) * (size_t)n, hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipStreamSynchronize(s) == hipSuccess;
    };
    auto fail_p = [&](int p, int code, const char* msg) {
        rcs[(size_t)p] = fail(ctxs[p], code, msg);
        ps[(size_t)p].running = false;
    };
    // Rounds of 8 iterations.  Each problem moves on to its next refinement step as soon as its
    // inner solve has converged (no barrier between problems): a round first issues the step
    // transitions of the problems that need one (fold the inner solution into phi64, new FP64
    // residual and inner start, first V-cycle), then 8 iterations of every problem; a problem
    // with nothing to do returns at once in those launches (its stop flag is set).
    static const double kStop = 1.0;
    auto stop = [&](int p) {  // a problem that leaves the rounds: its launches return at once
        return hipMemcpyAsync(&ctxs[p]->sc->conv, &kStop, sizeof kStop, hipMemcpyHostToDevice, s) == hipSuccess;
    };
    Plan iter8;
    {
        auto all = members([](const PS& q) { return q.running; });
        if (!all.empty() && !record_plan(all, mem, iter8, [&](int p) {
                for (uint32_t k = 0; k < 8; k++) iterate_once(ctxs[p], ps[(size_t)p].S);
            }))
            return MHDP_NO_MEMORY;
    }
    PlanGraph rounds;  // a capture or instantiation that fails leaves the rounds issued launch by launch
    if (batch_graph_enabled() && !iter8.ops.empty() && !rounds.build(iter8)) rounds.exec = nullptr;
    std::vector<int> trans = members([](const PS& q) { return q.running; });
    for (int p : trans) ps[(size_t)p].inner = true;
    for (bool first_round = true;; first_round = false) {
        if (!trans.empty()) {
            Plan tp;
            if (!record_plan(trans, mem, tp, [&](int p) {
                    mhdp_ctx* c = ctxs[p];
                    PS& q = ps[(size_t)p];
                    if (!first_round)
                        launch_outer_update(s, c->g, c->x64, c->x, c->p, c->p2, c->dinv, c->sc, q.S.meanfree,
                                            q.outer == 1);
                    if (!(first_round && q.started))
                        outer_start(c, s, q.S.meanfree, q.outer == 0);
                    mg_initialize(c, q.S);
                }))
                return MHDP_NO_MEMORY;
            issue_plan(s, tp);
            for (int p : trans) {
                ps[(size_t)p].it = 0;
                ps[(size_t)p].best_it = 0;
                ps[(size_t)p].best = std::numeric_limits<double>::infinity();
                ps[(size_t)p].checked = false;
                ps[(size_t)p].inner = true;
            }
        }
        auto now = members([](const PS& q) { return q.running; });
        if (now.empty()) break;
        if (rounds.exec) {
            if (hipGraphLaunch(rounds.exec, s) != hipSuccess) return MHDP_DEVICE;
        } else {
            issue_plan(s, iter8);
        }
        if (!read_all()) return MHDP_DEVICE;
        trans.clear();
        for (int p : now) {
            PS& q = ps[(size_t)p];
            const Scal& v = sh[(size_t)p];
            if (v.fault) {  // multigrid PCG breakdown: solved again on its own after the batch
                q.breakdown = true;
                q.running = false;
                continue;  // its stop flag (conv) is already set
            }
            const double vals[10] = {v.rr64, v.bb64, v.rz, v.rr, v.bb, v.sr, v.alpha, v.beta, v.conv, v.iters};
            bool bad = false;
            for (double x : vals) bad |= !std::isfinite(x);
            if (bad) {
                fail_p(p, MHDP_NONFINITE, "Nonfinite fluid pressure residual");
                if (!stop(p)) return MHDP_DEVICE;
                continue;
            }
            if (!q.checked) {
                q.res64 = v.bb64 > 0.0 ? std::sqrt(v.rr64 / v.bb64) : 0.0;
                q.checked = true;
                if (v.outer_done) {  // its stop flag is already set
                    q.done = true;
                    q.running = false;
                    continue;
                }
                if (q.outer >= max_outer) {
                    q.running = false;
                    if (!stop(p)) return MHDP_DEVICE;
                    continue;
                }
            }
            const bool conv = v.conv > 0.5;
            const uint32_t seen = (uint32_t)v.iters;
            if (seen <= q.it && !conv) {
                fail_p(p, MHDP_NO_PROGRESS, "Fluid pressure iteration made no progress");
                if (!stop(p)) return MHDP_DEVICE;
                continue;
            }
            q.it = seen;
            if (conv || q.it >= o->max_iter) {  // inner solve finished: next round moves it on
                if (!conv && !stop(p)) return MHDP_DEVICE;
                q.iterations += q.it;
                q.outer++;
                trans.push_back(p);
            } else if (v.rr < q.best) {
                q.best = v.rr;
                q.best_it = q.it;
            } else if (q.S.stall && q.it - q.best_it >= kStallIterations) {  // stalled: as a breakdown
                q.breakdown = true;
                q.running = false;
                if (!stop(p)) return MHDP_DEVICE;
            }
        }
    }
    // outputs of the problems that reached the target
    {
        auto fin = members([](const PS& q) { return q.done; });
        std::vector<char> copied((size_t)n, 0);  // the correction pass also wrote the potential
        Plan plan;
        if (!record_plan(fin, mem, plan, [&](int p) {
                mhdp_ctx* c = ctxs[p];
                if (ps[(size_t)p].outer == 0) launch_fill_f64(s, c->g.n, c->x64, 0.0);
                if (uout && uout[p])
                    copied[(size_t)p] = launch_correct_f64(s, c->g, c->grid.h, c->bx, c->by, c->bz, c->x64, ustar[p],
                                                           uout[p], phi ? phi[p] : nullptr);
            }))
            return MHDP_NO_MEMORY;
        issue_plan(s, plan);
        for (int p : fin)
            if (phi && phi[p] && !copied[(size_t)p] &&
                hipMemcpyAsync(phi[p], ctxs[p]->x64, (size_t)ctxs[p]->g.n * 8, hipMemcpyDeviceToDevice, s) != hipSuccess)
                return MHDP_DEVICE;
    }
    if (hipEventRecord(e1, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return MHDP_DEVICE;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    for (int p = 0; p < n; p++) {
        mhdp_ctx* c = ctxs[p];
        PS& q = ps[(size_t)p];
        if (rcs[(size_t)p] != MHDP_OK) continue;
        if (q.breakdown) {  // solved again on its own with the reference's V-cycle (its single projection
                            // would break down the same way: its results are those of the batch)
            c->timer.kernel = timers.kernel[(size_t)p];
            mhdp_stats* sp = st ? st + p : nullptr;
            int rc = project_once(c, ustar[p], o, kReferenceCycle, uout ? uout[p] : nullptr, phi ? phi[p] : nullptr, sp);
            if (rc == MHDP_OK && sp) sp->cycle_fallback = 1;
            if (rc == kBreakdown)
                rc = fail(c, MHDP_NO_PROGRESS,
                          "Fluid pressure iteration made no progress (the preconditioner is not positive definite)");
            rcs[(size_t)p] = rc;
            continue;
        }
        if (!q.done) {
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "Fluid pressure solve did not converge after %u iterations (true relative residual %.3e, "
                          "requested %.3e)",
                          q.iterations, q.res64, o->refine_tol);
            rcs[(size_t)p] = fail(c, MHDP_NOT_CONVERGED, buf);
            continue;
        }
        std::swap(c->x64, c->phi_prev);  // kept for a later warm start
        c->prev_valid = true;
        if (st) {
            mhdp_stats& t = st[p];
            t.iterations = q.iterations;
            t.outer_steps = q.outer;
            t.components = c->components;
            t.fluid_cells = c->fluid_cells;
            t.levels = (uint32_t)c->levels.size();
            t.meanfree = q.S.meanfree ? 1 : 0;
            t.multigrid = 1;
            t.rel_residual = (float)q.res64;
            t.rel_residual_f64 = q.res64;
            t.seconds = ms * 1e-3;
            t.cycle_fallback = 0;
        }
    }
    return hipGetLastError() == hipSuccess ? MHDP_OK : MHDP_DEVICE;
}

// MHDP_HOST_ANCHORS=1: label the gauge mask with the host BFS (comparison runs).
bool host_anchors_env() {
    const char* e = std::getenv("MHDP_HOST_ANCHORS");
    return e && e[0] == '1';
}

// MHDP_BATCH_MODE=threads: mhdp_project_batch drives per-context streams from host threads.
bool batch_threads_env() {
    const char* e = std::getenv("MHDP_BATCH_MODE");
    return e && std::strcmp(e, "threads") == 0;
}

// Host threads of mhdp_project_batch (MHDP_BATCH_THREADS, 1..64; default 8), read per call.
int batch_workers() {
    const char* e = std::getenv("MHDP_BATCH_THREADS");
    const int v = e ? std::atoi(e) : 0;
    return (v >= 1 && v <= 64) ? v : 8;
}

}  // namespace

// ============================================================================ static fields scratch
// The partial records of mhdp_fields_noalloc_dev live in a file-scope device array: one record of
// kFieldTotals doubles per block of the largest launch a tuning allows, then the folded totals
// (328 KB).  Each GPU holds its own copy for the life of the process; every call on a GPU shares
// it (every block writes its whole record before the fold reads it, so nothing needs clearing), and
// calls that overlap on one GPU race.  Host code never passes the array's name to a kernel: that is
// a host-side stand-in, not a device address.  The address is taken once per device with
// hipGetSymbolAddress, with that device current, and kept.
namespace {
constexpr size_t kFieldsScratch = ((size_t)kMaxPartials + 1) * kFieldTotals;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) double scratch[kFieldsScratch];

// This device's address of `scratch` (nullptr when the device cannot be named or the call fails).
static double* fields_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}

// ============================================================================ launch tuning
namespace {
uint32_t env_u32(const char* name, uint32_t lo, uint32_t hi, uint32_t def) {
    const char* e = std::getenv(name);
    const long v = e ? std::strtol(e, nullptr, 10) : 0;
    return (v >= (long)lo && v <= (long)hi) ? (uint32_t)v : def;
}
bool tuning_valid(const mhdp_launch_tuning& t) {
    // kMaxPartials bounds every per-block partial buffer
    auto cap = [](uint32_t v) { return v >= 8 && v <= kMaxPartials; };
    return cap(t.grid_cap) && cap(t.brick_cap) && cap(t.march_cap) && t.march_z >= 2 && t.march_z <= 512 &&
           t.march_z % 2 == 0 && t.brick_occ >= 2 && t.brick_occ <= 4;
}
}  // namespace

LaunchTuning& launch_tuning() {
    static LaunchTuning t = [] {
        LaunchTuning d;
        d.grid_cap = env_u32("MHDP_GRID_CAP", 8, 4096, 1024);
        d.brick_cap = env_u32("MHDP_BRICK_CAP", 8, 4096, 1024);
        d.march_cap = env_u32("MHDP_MARCH_CAP", 8, 4096, 1024);
        const uint32_t z = env_u32("MHDP_MARCH_Z", 2, 512, 32);
        d.march_z = z % 2 == 0 ? z : 32;
        d.brick_occ = (int)env_u32("MHDP_BRICK_OCC", 2, 4, 2);
        return d;
    }();
    return t;
}

// ============================================================================ C ABI

extern "C" {

int mhdp_get_launch_tuning(mhdp_launch_tuning* out) {
    if (!out) return MHDP_INVALID;
    const LaunchTuning& t = launch_tuning();
    *out = mhdp_launch_tuning{t.grid_cap, t.brick_cap, t.march_cap, t.march_z, t.brick_occ};
    return MHDP_OK;
}

int mhdp_set_launch_tuning(const mhdp_launch_tuning* in) {
    if (!in || !tuning_valid(*in)) return MHDP_INVALID;
    LaunchTuning& t = launch_tuning();
    t.grid_cap = in->grid_cap;
    t.brick_cap = in->brick_cap;
    t.march_cap = in->march_cap;
    t.march_z = in->march_z;
    t.brick_occ = in->brick_occ;
    return MHDP_OK;
}

int mhdp_create(mhdp_ctx** out, const mhdp_grid* grid, const mhdp_phase* phase) {
    if (!out || !grid || !phase) return MHDP_INVALID;
    *out = nullptr;
    const uint64_t n = (uint64_t)grid->nx * grid->ny * grid->nz;
    if (grid->nx == 0 || grid->ny == 0 || grid->nz == 0 || n > (1ull << 30) || face_count(*grid) > 0xffffffffull ||
        !(grid->h > 0.0f) || !std::isfinite(grid->h) || grid->periodic > 7u)
        return MHDP_INVALID;
    if (phase->two_phase && (!(phase->rho_liquid > 0.0f) || !(phase->rho_air > 0.0f) ||
                             !std::isfinite(phase->rho_liquid) || !std::isfinite(phase->rho_air)))
        return MHDP_INVALID;
    mhdp_ctx* c = new mhdp_ctx();
    c->grid = *grid;
    c->phase = *phase;
    if (!phase->two_phase) c->phase.rho_liquid = phase->rho_liquid > 0.0f ? phase->rho_liquid : 1.0f;
    c->g.nx = (int)grid->nx;
    c->g.ny = (int)grid->ny;
    c->g.nz = (int)grid->nz;
    c->g.per = grid->periodic;
    c->g.n = (uint32_t)n;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc((void**)&c->sc_h, sizeof(Scal), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->fields_h, kFieldTotals * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        mhdp_destroy(c);
        return MHDP_DEVICE;
    }
    c->own_stream = true;
    if (const char* e = std::getenv("MHDP_UNFUSED")) c->fused_enabled = e[0] != '1';
    c->brick_enabled = brick_enabled_env();
    int rc = ensure_fine(c);
    if (rc) {
        mhdp_destroy(c);
        return rc;
    }
    // the reduction tickets start (and always end) at zero
    std::memset(c->sc_h, 0, sizeof(Scal));
    if (hipMemcpy(c->sc, c->sc_h, sizeof(Scal), hipMemcpyHostToDevice) != hipSuccess) {
        mhdp_destroy(c);
        return MHDP_DEVICE;
    }
    *out = c;
    return MHDP_OK;
}

int mhdp_set_stream(mhdp_ctx* c, void* stream) {
    if (!c) return MHDP_INVALID;
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
    return MHDP_OK;
}

int mhdp_set_cells(mhdp_ctx* c, const uint8_t* fluid, const float* alpha) {
    if (!c) return MHDP_INVALID;
    if (!fluid) return fail(c, MHDP_INVALID, "Missing fluid mask");
    if (c->phase.two_phase && !alpha) return fail(c, MHDP_INVALID, "Two-phase projection needs the liquid fraction");
    const uint32_t n = c->g.n;
    if (alpha)
        for (uint32_t i = 0; i < n; i++)
            if (!std::isfinite(alpha[i])) return fail(c, MHDP_INVALID, "Nonfinite liquid fraction");
    c->cells = false;
    c->open8_fresh = false;
    int rc;
    if ((rc = dalloc(c, (void**)&c->fluid8, n)) || (rc = dalloc(c, (void**)&c->cc_counts, 4 * sizeof(uint32_t))))
        return rc;
    HIP_OK(hipMemcpyAsync(c->fluid8, fluid, n, hipMemcpyHostToDevice, c->stream));
    if (alpha) HIP_OK(hipMemcpyAsync(c->alpha, alpha, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    // anchors() on the device; the host BFS takes over if a union-find bound was reached
    // (never seen; the bounds exist so that no input can keep the kernels running).
    uint32_t counts[4] = {0u, 0u, 0u, 0u};
    bool host = host_anchors_env();
    if (!host) {
        launch_components(c->stream, c->g, c->fluid8, c->mask, c->cc_counts);
        HIP_OK(hipMemcpyAsync(counts, c->cc_counts, sizeof counts, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        HIP_OK(hipGetLastError());
        host = counts[3] != 0u;
    }
    if (host) {
        std::vector<uint32_t> mask(n);
        uint32_t fluid_cells = 0;
        for (uint32_t i = 0; i < n; i++) {
            mask[i] = fluid[i] ? 1u : 0u;
            fluid_cells += mask[i];
        }
        const uint32_t d[3] = {(uint32_t)c->g.nx, (uint32_t)c->g.ny, (uint32_t)c->g.nz};
        counts[0] = label_components(mask, d, c->g.per);
        counts[1] = fluid_cells;
        counts[2] = 0u;
        for (uint32_t i = 0; i < n; i++)
            if ((mask[i] & 3u) == 2u) {
                counts[2] = i;
                break;
            }
        HIP_OK(hipMemcpyAsync(c->mask, mask.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    c->has_alpha = alpha != nullptr;
    c->components = counts[0];
    c->fluid_cells = counts[1];
    c->anchor0 = counts[0] ? counts[2] : 0u;
    c->labels_on_host = host;
    // the per-cell open-face bytes of the step and fields kernels depend only on the mask: built
    // here, with the set-up, so no later entry point has to allocate them
    if ((rc = dalloc(c, (void**)&c->open8, n))) return rc;
    launch_open8(c->stream, c->g, c->mask, c->open8);
    HIP_OK(hipStreamSynchronize(c->stream));
    c->open8_fresh = true;
    c->cells = true;
    return MHDP_OK;
}

int mhdp_get_mask(const mhdp_ctx* c, uint32_t* encoded) {
    if (!c || !encoded || !c->cells) return MHDP_INVALID;
    if (hipMemcpyAsync(encoded, c->mask, (size_t)c->g.n * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return MHDP_DEVICE;
    return MHDP_OK;
}

int mhdp_set_potential(mhdp_ctx* c, const double* phi, int on_device) {
    if (!c || !phi) return MHDP_INVALID;
    const size_t bytes = (size_t)c->g.n * 8;
    int rc = dalloc(c, (void**)&c->phi_prev, bytes);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(c->phi_prev, phi, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->prev_valid = true;
    return MHDP_OK;
}

// Argument checks shared by the fields entry points.
static int fields_check(mhdp_ctx* c, const float* u_faces, const double* phi, const mhdp_fields_params* fp,
                        const float* cells) {
    if (!c->cells) return fail(c, MHDP_INVALID, "Fluid cells are not set");
    if (!u_faces || !phi || !fp || !cells) return fail(c, MHDP_INVALID, "Missing fields argument");
    if ((reinterpret_cast<uintptr_t>(cells) & 15u) != 0)
        return fail(c, MHDP_INVALID, "Cell packets must be 16-byte aligned");
    if (!(fp->dt > 0.0f) || !std::isfinite(fp->dt) || (fp->gauge != MHDP_GAUGE_AUTO && fp->gauge != MHDP_GAUGE_PINNED))
        return fail(c, MHDP_INVALID, "Invalid fields parameters");
    if (face_count(c->grid) >= (1ull << 32)) return fail(c, MHDP_INVALID, "Too many faces for the step kernels");
    return MHDP_OK;
}

// The fields + diagnostics launches and read_diagnostics, given the device partial records `part`
// ((blocks + 1) * kFieldTotals doubles: one record per block, then the folded totals) and a host
// buffer of kFieldTotals doubles for the totals.  Allocates nothing.
static int fields_run(mhdp_ctx* c, const float* u_faces, const double* phi, const float* em,
                      const mhdp_fields_params* fp, float* cells, mhdp_diagnostics* diag, double* part, double* t) {
    const uint32_t blocks = grid_blocks(c->g.n);
    const bool meanfree = fp->gauge != MHDP_GAUGE_PINNED && c->components == 1;
    double* totals = part + (size_t)blocks * kFieldTotals;
    launch_fields(c->stream, c->g, c->grid.h, c->phase.rho_liquid, c->phase.rho_air, c->phase.two_phase ? 1 : 0,
                  fp->dt, fp->acceleration, fp->mushy ? 1 : 0, meanfree ? 1 : 0, c->mask, c->open8, u_faces, phi, em,
                  cells, part, totals);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(t, totals, kFieldTotals * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    // read_diagnostics (src/fluid/mod.rs:1116-1146)
    const double h3 = (double)c->grid.h * c->grid.h * c->grid.h;
    mhdp_diagnostics d{};
    d.max_speed = t[0];
    d.max_divergence = t[1];
    d.rms_divergence = std::sqrt(t[2] / (double)std::max<uint32_t>(c->fluid_cells, 1u));
    d.kinetic_energy = t[3];
    d.max_pressure = t[4];
    d.net_flux = t[5] * h3;
    d.max_force = t[6];
    d.nonfinite_cells = (uint32_t)t[7];
    d.max_face_speed = t[8];
    d.fluid_mass = (double)c->phase.rho_liquid * t[9] * h3;
    if (diag) *diag = d;
    for (int k = 0; k < kFieldTotals; k++)
        if (!std::isfinite(t[k])) return fail(c, MHDP_NONFINITE, "Nonfinite fluid fields; state restored");
    if (t[7] != 0.0) return fail(c, MHDP_NONFINITE, "Nonfinite fluid fields; state restored");
    return MHDP_OK;
}

int mhdp_fields_dev(mhdp_ctx* c, const float* u_faces, const double* phi, const float* em,
                    const mhdp_fields_params* fp, float* cells, mhdp_diagnostics* diag) {
    if (!c) return MHDP_INVALID;
    c->err.clear();
    int rc = fields_check(c, u_faces, phi, fp, cells);
    if (rc) return rc;
    // sized for the largest block cap a launch tuning allows (the buffer is kept across calls)
    if ((rc = dalloc(c, (void**)&c->fields_part, ((size_t)kMaxPartials + 1) * kFieldTotals * sizeof(double))))
        return rc;
    if (!c->open8_fresh) {
        if ((rc = dalloc(c, (void**)&c->open8, c->g.n))) return rc;
        launch_open8(c->stream, c->g, c->mask, c->open8);
        c->open8_fresh = true;
    }
    double t[kFieldTotals];
    return fields_run(c, u_faces, phi, em, fp, cells, diag, c->fields_part, t);
}

int mhdp_fields_noalloc_dev(mhdp_ctx* c, const float* u_faces, const double* phi, const float* em,
                            const mhdp_fields_params* fp, float* cells, mhdp_diagnostics* diag) {
    if (!c) return MHDP_INVALID;
    c->err.clear();
    int rc = fields_check(c, u_faces, phi, fp, cells);
    if (rc) return rc;
    if (!c->open8_fresh || !c->open8) return fail(c, MHDP_INVALID, "Fluid cells are not set");
    const uint32_t blocks = grid_blocks(c->g.n);
    if (((size_t)blocks + 1) * kFieldTotals > kFieldsScratch)
        return fail(c, MHDP_INVALID, "Fields partial records exceed the static scratch (" +
                                         std::to_string(kFieldsScratch / kFieldTotals - 1) + " blocks)");
    double* part = fields_scratch();
    if (!part) return fail(c, MHDP_DEVICE, "HIP error: static fields scratch has no device address");
    return fields_run(c, u_faces, phi, em, fp, cells, diag, part, c->fields_h);
}

static int step_check(mhdp_ctx* c, const mhdp_step_params* sp) {
    if (!c->cells) return fail(c, MHDP_INVALID, "Fluid cells are not set");
    if (!sp || !(sp->dt > 0.0f) || !std::isfinite(sp->dt) || !std::isfinite(sp->viscosity) || sp->viscosity < 0.0f ||
        !std::isfinite(sp->surface_tension) || !std::isfinite(sp->acceleration[0]) ||
        !std::isfinite(sp->acceleration[1]) || !std::isfinite(sp->acceleration[2]))
        return fail(c, MHDP_INVALID, "Invalid step parameters");
    if (face_count(c->grid) >= (1ull << 32)) return fail(c, MHDP_INVALID, "Too many faces for the step kernels");
    if (!c->open8_fresh) {
        int rc = dalloc(c, (void**)&c->open8, c->g.n);
        if (rc) return rc;
        launch_open8(c->stream, c->g, c->mask, c->open8);
        c->open8_fresh = true;
    }
    return MHDP_OK;
}

// RKC2 stage rows of the damped Runge-Kutta-Chebyshev method (Verwer, Hundsdorfer & Sommeijer
// 1990, damping 0.05) as the reference schedules them (src/fluid/stabilized.rs:13-53): the fewest
// stages whose stability interval covers 1.8 x euler_steps; rows of (mu, nu, 1 - mu - nu,
// operator weight, initial-operator weight).
static std::vector<std::array<float, 5>> rkc_schedule(double dt, uint32_t euler_steps) {
    const double damping = 0.05, stiffness = 1.8 * (double)euler_steps;
    const int capacity = (int)std::ceil(3.0 * std::sqrt((double)euler_steps)) + 2;
    for (int s = 2; s <= capacity; s++) {
        const double w0 = 1.0 + damping / ((double)s * s);
        std::vector<double> t(s + 1), d(s + 1), dd(s + 1), b(s + 1), a(s + 1);
        t[0] = 1.0; t[1] = w0; d[0] = 0.0; d[1] = 1.0; dd[0] = 0.0; dd[1] = 0.0;
        for (int j = 2; j <= s; j++) {
            t[j] = 2.0 * w0 * t[j - 1] - t[j - 2];
            d[j] = 2.0 * t[j - 1] + 2.0 * w0 * d[j - 1] - d[j - 2];
            dd[j] = 4.0 * d[j - 1] + 2.0 * w0 * dd[j - 1] - dd[j - 2];
        }
        const double w1 = d[s] / dd[s];
        if (stiffness > 0.9 * (1.0 + w0) / w1) continue;
        for (int j = 2; j <= s; j++) b[j] = dd[j] / (d[j] * d[j]);
        b[0] = b[2];
        b[1] = b[2];
        for (int j = 0; j <= s; j++) a[j] = 1.0 - b[j] * t[j];
        std::vector<std::array<float, 5>> rows;
        rows.push_back({1.0f, 0.0f, 0.0f, (float)(dt * b[1] * w1), 0.0f});
        for (int j = 2; j <= s; j++) {
            const float mu = (float)(2.0 * w0 * b[j] / b[j - 1]);
            const float nu = (float)(-b[j] / b[j - 2]);
            const double op = dt * 2.0 * w1 * b[j] / b[j - 1];
            rows.push_back({mu, nu, 1.0f - mu - nu, (float)op, (float)(-a[j - 1] * op)});
        }
        return rows;
    }
    return {};
}

int mhdp_surface_stress_dev(mhdp_ctx* c, float* u, const float* cells, const mhdp_stress_params* sp, uint32_t* substeps) {
    if (!c) return MHDP_INVALID;
    c->err.clear();
    if (substeps) *substeps = 0;
    if (!c->cells) return fail(c, MHDP_INVALID, "Fluid cells are not set");
    if (!c->phase.two_phase) return fail(c, MHDP_INVALID, "Surface stress diffusion needs a two-phase context");
    if (!u || !cells || !sp) return fail(c, MHDP_INVALID, "Missing surface stress argument");
    if (!(sp->dt > 0.0) || !std::isfinite(sp->dt) || !(sp->viscosity >= 0.0f) || !std::isfinite(sp->viscosity) ||
        !(sp->mu_air >= 0.0f) || !std::isfinite(sp->mu_air) || sp->max_substeps < 1 || sp->max_substeps > 4096)
        return fail(c, MHDP_INVALID, "Invalid surface stress parameters");
    if (face_count(c->grid) >= (1ull << 32)) return fail(c, MHDP_INVALID, "Too many faces for the step kernels");
    int rc;
    if (!c->open8_fresh) {
        if ((rc = dalloc(c, (void**)&c->open8, c->g.n))) return rc;
        launch_open8(c->stream, c->g, c->mask, c->open8);
        c->open8_fresh = true;
    }
    const size_t F = face_count(c->grid);
    if ((rc = dalloc(c, (void**)&c->stress_mu, (size_t)c->g.n * 4)) ||
        (rc = dalloc(c, (void**)&c->stress_rho, (size_t)c->g.n * 4)) ||
        (rc = dalloc(c, (void**)&c->stress_part, ((size_t)stress_bound_blocks(c->g) + 1) * 4)) ||
        (rc = dalloc(c, (void**)&c->stress_tensor, stress_tensor_len(c->g) * 4)))
        return rc;
    const float h = c->grid.h, rl = c->phase.rho_liquid, ra = c->phase.rho_air;
    const float nu = sp->viscosity, mua = sp->mu_air;
    const double dt = sp->dt;
    // global_viscous_bound (mod.rs:1146-1158): one substep is proved safe without a reduction
    const double eps = 1.1920928955078125e-07;
    const double slack = (double)1e-5f + 4.0 * eps;  // surface::FRACTION_TOLERANCE + 4 f32 eps
    const double floor_rho = std::min((double)rl, (double)ra) - std::fabs((double)rl - (double)ra) * slack;
    double count = 0.0;
    bool proved = false;
    if (floor_rho > 0.0) {
        const double mu = (double)std::max(nu * rl, mua);
        const double gb = 24.0 * mu / (floor_rho * (double)h * (double)h) * (1.0 + 32.0 * eps);
        if (dt * gb <= 1.8) {
            if (gb == 0.0) return MHDP_OK;
            count = 1.0;
            proved = true;
        }
    }
    launch_stress_cells(c->stream, c->g, c->mask, cells, nu, mua, rl, ra, c->stress_mu, c->stress_rho);
    const StressBufs sb{h, nu, mua, rl, ra, c->mask, c->open8, c->stress_mu, c->stress_rho, c->stress_tensor};
    if (!proved) {
        const uint32_t nb = stress_bound_blocks(c->g);
        launch_stress_bound(c->stream, c->g, sb, c->stress_part, c->stress_part + nb);
        float bound = 0.0f;
        HIP_OK(hipMemcpyAsync(&bound, c->stress_part + nb, 4, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        if (!std::isfinite(bound) || bound < 0.0f) return fail(c, MHDP_NONFINITE, "Nonfinite viscous stability bound");
        if (bound == 0.0f) return MHDP_OK;
        count = std::max(std::ceil(dt * (double)bound / 1.8), 1.0);
    }
    if (count > (double)sp->max_substeps) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "Viscosity requires %.0f equivalent Euler substeps, above the %u work limit; reduce the common "
                      "timestep or increase Viscosity work limit",
                      count, sp->max_substeps);
        return fail(c, MHDP_INVALID, buf);
    }
    const uint32_t n = (uint32_t)count;
    if (n > 8) {  // RKC2 (stabilized.rs:199-250)
        const auto rows = rkc_schedule(dt, n);
        if (rows.empty()) return fail(c, MHDP_INVALID, "RKC stage-capacity bound");
        if ((rc = dalloc(c, (void**)&c->faces_b, F * 4)) || (rc = dalloc(c, (void**)&c->faces_y0, F * 4)) ||
            (rc = dalloc(c, (void**)&c->faces_f0, F * 4)))
            return rc;
        HIP_OK(hipMemcpyAsync(c->faces_y0, u, F * 4, hipMemcpyDeviceToDevice, c->stream));
        launch_stress_tensor(c->stream, c->g, sb, 1, u);
        launch_stress_initial(c->stream, c->g, sb, u, c->faces_f0);
        HIP_OK(hipMemsetAsync(c->faces_b, 0, F * 4, c->stream));
        float* newest = u;
        float* other = c->faces_b;
        for (size_t i = 0; i < rows.size(); i++) {
            if (i > 0) launch_stress_tensor(c->stream, c->g, sb, 1, newest);  // stage 0 reads u's tensor
            launch_stress_stage(c->stream, c->g, sb, rows[i].data(), newest, other, c->faces_y0, c->faces_f0, other);
            std::swap(newest, other);
        }
        if (newest != u) HIP_OK(hipMemcpyAsync(u, newest, F * 4, hipMemcpyDeviceToDevice, c->stream));
        if (substeps) *substeps = (uint32_t)rows.size();
    } else {  // explicit substeps of dt / count (mod.rs:1219-1233), in place: each substep's tensor
              // holds every neighbour value its face kernels need
        const float dt_sub = (float)(dt / (double)n);
        for (uint32_t k = 0; k < n; k++) {
            launch_stress_tensor(c->stream, c->g, sb, 0, u);
            launch_stress_euler(c->stream, c->g, sb, dt_sub, u, u);
        }
        if (substeps) *substeps = n;
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

int mhdp_predict_dev(mhdp_ctx* c, const float* u, const mhdp_step_params* sp, float* out) {
    if (!c) return MHDP_INVALID;
    c->err.clear();
    int rc = step_check(c, sp);
    if (rc) return rc;
    if (!u || !out) return fail(c, MHDP_INVALID, "Missing face velocities");
    if ((rc = dalloc(c, (void**)&c->faces_open, face_count(c->grid) * sizeof(float)))) return rc;
    launch_predict(c->stream, c->g, c->grid.h, sp->dt, sp->advect ? 1 : 0, c->open8, u, c->faces_open, out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

int mhdp_viscous_dev(mhdp_ctx* c, const float* pred, const float* cells, const float* em, const mhdp_step_params* sp,
                     float* out) {
    if (!c) return MHDP_INVALID;
    c->err.clear();
    int rc = step_check(c, sp);
    if (rc) return rc;
    if (!pred || !out) return fail(c, MHDP_INVALID, "Missing face velocities");
    if (pred == out) return fail(c, MHDP_INVALID, "viscous reads neighbouring faces: output must not alias input");
    if (c->phase.two_phase && !cells) return fail(c, MHDP_INVALID, "Two-phase viscous step needs the cell packets");
    if (c->phase.two_phase && (rc = dalloc(c, (void**)&c->cell_pair, (size_t)c->g.n * 8))) return rc;
    launch_viscous(c->stream, c->g, c->grid.h, c->phase.rho_liquid, c->phase.rho_air, c->phase.two_phase ? 1 : 0,
                   sp->dt, sp->viscosity, sp->acceleration, sp->surface_tension, c->open8, pred, cells,
                   (float2*)c->cell_pair, em, out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

int mhdp_project_dev(mhdp_ctx* c, const float* ustar, const mhdp_opts* o, float* uout, double* phi, mhdp_stats* st) {
    if (!c) return MHDP_INVALID;
    c->err.clear();
    return project_impl(c, ustar, o, uout, phi, st);
}

int mhdp_project(mhdp_ctx* c, const float* ustar, const mhdp_opts* o, float* uout, double* phi, mhdp_stats* st) {
    if (!c) return MHDP_INVALID;
    c->err.clear();
    if (!ustar) return fail(c, MHDP_INVALID, "Missing candidate face velocities");
    const size_t fb = face_count(c->grid) * sizeof(float);
    int rc;
    if ((rc = dalloc(c, (void**)&c->ustar_stage, fb)) || (rc = dalloc(c, (void**)&c->uout_stage, fb)) ||
        (rc = dalloc(c, (void**)&c->phi_stage, (size_t)c->g.n * 8)))
        return rc;
    HIP_OK(hipMemcpyAsync(c->ustar_stage, ustar, fb, hipMemcpyHostToDevice, c->stream));
    rc = project_impl(c, c->ustar_stage, o, c->uout_stage, c->phi_stage, st);
    if (rc) return rc;
    if (uout) HIP_OK(hipMemcpyAsync(uout, c->uout_stage, fb, hipMemcpyDeviceToHost, c->stream));
    if (phi) HIP_OK(hipMemcpyAsync(phi, c->phi_stage, (size_t)c->g.n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

int mhdp_project_batch(mhdp_ctx* const* ctxs, int count, const float* const* ustar, const mhdp_opts* o,
                       float* const* uout, double* const* phi, mhdp_stats* st) {
    if (!ctxs || count < 0 || !ustar || !o) return MHDP_INVALID;
    std::vector<const mhdp_ctx*> seen(ctxs, ctxs + count);
    std::sort(seen.begin(), seen.end());
    if (count && (!seen[0] || std::adjacent_find(seen.begin(), seen.end()) != seen.end())) return MHDP_INVALID;
    std::vector<int> rcs((size_t)count, MHDP_OK);
    // The fused multigrid path of the mixed solver runs in lockstep: one launch per kernel for
    // all problems (MHDP_BATCH_MODE=threads keeps the per-context streams below).
    bool lockstep = count > 0 && !batch_threads_env();
    for (int k = 0; k < count && lockstep; k++) lockstep = lockstep_ok(ctxs[k], o);
    if (lockstep) {
        const int rc = project_lockstep(ctxs, count, ustar, o, uout, phi, st, rcs);
        if (rc) return rc;
        for (int k = 0; k < count; k++)
            if (rcs[(size_t)k]) return rcs[(size_t)k];
        return MHDP_OK;
    }
    // Otherwise each context runs on its own stream; a small pool of host threads drives them so
    // the problems' kernels and host synchronisations overlap.  Largest grids start first.
    std::vector<int> order((size_t)count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ctxs[a]->g.n > ctxs[b]->g.n; });
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return MHDP_DEVICE;
    std::atomic<int> next{0};
    auto work = [&] {
        if (hipSetDevice(device) != hipSuccess) return;
        for (;;) {
            const int j = next.fetch_add(1);
            if (j >= count) break;
            const int k = order[(size_t)j];
            mhdp_ctx* c = ctxs[k];
            c->err.clear();
            rcs[(size_t)k] = project_impl(c, ustar[k], o, uout ? uout[k] : nullptr, phi ? phi[k] : nullptr,
                                          st ? st + k : nullptr);
        }
    };
    const int workers = std::min(count, batch_workers());
    if (workers <= 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        pool.reserve((size_t)workers);
        for (int w = 0; w < workers; w++) pool.emplace_back(work);
        for (auto& t : pool) t.join();
    }
    if (next.load() < count) return MHDP_DEVICE;  // a worker could not select the device
    for (int k = 0; k < count; k++)
        if (rcs[(size_t)k]) return rcs[(size_t)k];
    return MHDP_OK;
}

uint32_t mhdp_anchors(uint32_t* mask, const uint32_t dims[3], uint32_t periodic) {
    if (!mask || !dims) return 0;
    const uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
    std::vector<uint32_t> m(mask, mask + n);
    for (auto& v : m) v = v ? 1u : 0u;
    const uint32_t comps = label_components(m, dims, periodic);
    std::memcpy(mask, m.data(), n * 4);
    return comps;
}

int mhdp_hierarchy(const uint32_t dims4[4], uint32_t* level_dims, int max_levels, uint64_t* coarse_cells) {
    if (!dims4) return -1;
    uint64_t count = 1;
    for (int a = 0; a < 3; a++) {
        if (dims4[a] == 0 || dims4[a] > 0x7fffffffu) return -1;
        count *= dims4[a];
        if (count > 0xffffffffull) return -1;
    }
    if (count != dims4[3]) return -1;
    uint32_t cur[4] = {dims4[0], dims4[1], dims4[2], dims4[3]};
    int levels = 0;
    uint64_t offset = 0;
    for (;;) {
        if (level_dims && levels < max_levels) std::memcpy(level_dims + 4 * levels, cur, sizeof cur);
        levels++;
        if (cur[0] <= 4 && cur[1] <= 4 && cur[2] <= 4) break;
        for (int a = 0; a < 3; a++) cur[a] = (cur[a] + 1) / 2;
        cur[3] = cur[0] * cur[1] * cur[2];
        offset += cur[3];
        if (offset > 0xffffffffull) return -1;
    }
    if (coarse_cells) *coarse_cells = offset;
    return levels;
}

int mhdp_set_kernel_timing(mhdp_ctx* c, int kernel) {
    if (!c || kernel < MHDP_K_NONE || kernel > MHDP_K_RESTRICT) return MHDP_INVALID;
    c->timer.kernel = kernel;
    return MHDP_OK;
}

int mhdp_kernel_time(const mhdp_ctx* c, double* total_ms, uint32_t* launches) {
    if (!c) return MHDP_INVALID;
    if (total_ms) *total_ms = c->timer_ms;
    if (launches) *launches = c->timer_launches;
    return MHDP_OK;
}

const char* mhdp_last_error(const mhdp_ctx* c) { return c ? c->err.c_str() : "invalid context"; }

uint64_t mhdp_device_bytes(const mhdp_ctx* c) { return c ? c->bytes : 0; }

void mhdp_destroy(mhdp_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    else (void)hipDeviceSynchronize();
    void* ptrs[] = {c->mask, c->alpha, c->bx, c->by, c->bz, c->diag, c->dinv, c->b, c->x, c->r, c->r2, c->p, c->p2, c->q, c->z,
                    c->t, c->x64, c->b64, c->bx16, c->by16, c->bz16, c->dinv16, c->sc, c->part, c->ustar_stage,
                    c->uout_stage, c->phi_stage, c->fluid8, c->cc_counts, c->phi_prev, c->uwarm, c->fields_part, c->faces_open, c->open8, c->stress_mu, c->stress_rho, c->stress_part, c->faces_b, c->faces_y0, c->faces_f0, c->stress_tensor, c->cell_pair};
    for (void* p : ptrs) dfree(c, p);
    for (CoarseArrays& a : c->coarse) {
        void* q[] = {a.bx, a.by, a.bz, a.react, a.diag, a.dinv, a.bx16, a.by16, a.bz16, a.react16, a.dinv16,
                     a.r, a.z, a.t, a.w};
        for (void* p : q) dfree(c, p);
    }
    for (hipEvent_t e : c->timer.ev) (void)hipEventDestroy(e);
    if (c->sc_h) (void)hipHostFree(c->sc_h);
    if (c->fields_h) (void)hipHostFree(c->fields_h);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// ---------------------------------------------------------------- test hooks

int mhdp_probe_coefficients(mhdp_ctx* c, int meanfree, float* beta3, float* diag, float* dinv) {
    if (!c) return MHDP_INVALID;
    int rc = prepare(c, meanfree != 0);
    if (rc) return rc;
    const size_t fb = (size_t)c->g.n * 4;
    if (beta3) {
        HIP_OK(hipMemcpyAsync(beta3, c->bx, fb, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(beta3 + c->g.n, c->by, fb, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(beta3 + 2 * (size_t)c->g.n, c->bz, fb, hipMemcpyDeviceToHost, c->stream));
    }
    if (diag) HIP_OK(hipMemcpyAsync(diag, c->diag, fb, hipMemcpyDeviceToHost, c->stream));
    if (dinv) HIP_OK(hipMemcpyAsync(dinv, c->dinv, fb, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

int mhdp_probe_apply(mhdp_ctx* c, int meanfree, const float* p, float* q) {
    if (!c || !p || !q) return MHDP_INVALID;
    int rc = prepare(c, meanfree != 0);
    if (rc) return rc;
    const size_t fb = (size_t)c->g.n * 4;
    HIP_OK(hipMemcpyAsync(c->p, p, fb, hipMemcpyHostToDevice, c->stream));
    launch_apply_plain(c->stream, c->g, c->p, c->bx, c->by, c->bz, c->dinv, 1, c->q);
    HIP_OK(hipMemcpyAsync(q, c->q, fb, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipGetLastError());
    return MHDP_OK;
}

int mhdp_probe_vcycle(mhdp_ctx* c, int meanfree, int precision, float smooth_weight, float coarse_scale,
                      int coarse_visits, int smoother, const float* r, float* z) {
    if (!c || !r || !z || coarse_visits < 0 || coarse_visits > kMaxVisits || smoother < MHDP_SMOOTH_AUTO ||
        smoother > MHDP_SMOOTH_RED_BLACK)
        return MHDP_INVALID;
    c->coarse_visits = coarse_visits_for(c, coarse_visits);
    c->rb = smoother == MHDP_SMOOTH_RED_BLACK;
    const bool f16 = precision == MHDP_MIXED;
    int rc = prepare(c, meanfree != 0);
    if (rc) return rc;
    if (!ensure_mg(c, f16)) return fail(c, MHDP_NO_MEMORY, "Optional pressure preconditioner is unavailable");
    build_mg(c, meanfree != 0, f16);
    Scal s;
    set_solve_params(c, s, meanfree != 0, 0.0, 0.0, 1.0, smooth_weight, coarse_scale);
    if ((rc = upload_scal(c, s))) return rc;
    const size_t fb = (size_t)c->g.n * 4;
    HIP_OK(hipMemcpyAsync(c->r, r, fb, hipMemcpyHostToDevice, c->stream));
    Solve S;
    S.meanfree = meanfree != 0;
    S.mg = true;
    S.store = f16 ? STORE_BF16 : STORE_F32;
    S.coarse64 = f16;
    S.lv = level_view(c);
    S.fused = S.lv.size() >= 2 && c->fused_enabled;
    void* zf = nullptr;
    if (S.fused) {  // the solver's own cycle: fine level fused into two passes
        vcycle_fused(c, S, FIN_NONE);
        zf = S.lv[0].t;
    } else {
        launch_vcycle(c->stream, S.store, (int)S.lv.size(), S.lv.data(), c->r, c->mask, c->sc, S.coarse64, c->coarse_visits,
                  &zf);
    }
    if (f16) {
        launch_from_bf16(c->stream, c->g.n, zf, c->q);
        HIP_OK(hipMemcpyAsync(z, c->q, fb, hipMemcpyDeviceToHost, c->stream));
    } else {
        HIP_OK(hipMemcpyAsync(z, zf, fb, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipGetLastError());
    return MHDP_OK;
}

int mhdp_probe_mg_graph(mhdp_ctx* c, int meanfree, int level, float* graph5, uint32_t* dims4) {
    if (!c) return MHDP_INVALID;
    int rc = prepare(c, meanfree != 0);
    if (rc) return rc;
    if (!ensure_mg(c, false)) return fail(c, MHDP_NO_MEMORY, "Optional pressure preconditioner is unavailable");
    if (level < 1 || level >= (int)c->levels.size()) return fail(c, MHDP_INVALID, "No such multigrid level");
    build_mg(c, meanfree != 0, false);
    const Grid3& g = c->levels[level];
    if (dims4) {
        dims4[0] = (uint32_t)g.nx;
        dims4[1] = (uint32_t)g.ny;
        dims4[2] = (uint32_t)g.nz;
        dims4[3] = g.n;
    }
    if (graph5) {
        CoarseArrays& a = c->coarse[level - 1];
        const size_t fb = (size_t)g.n * 4;
        const float* src[5] = {a.bx, a.by, a.bz, a.react, a.diag};
        for (int k = 0; k < 5; k++)
            HIP_OK(hipMemcpyAsync(graph5 + (size_t)k * g.n, src[k], fb, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

int mhdp_probe_multigrid(mhdp_ctx* c, int action, int arg, int* state, uint64_t* bytes) {
    if (!c || action < 0 || action > 2) return MHDP_INVALID;
    if (action == 1) c->mg_fail_after = arg < 0 ? 0 : arg;
    if (action == 2) {
        release_mg(c);
        c->mg_refused = true;
    }
    if (state) *state = c->mg_refused ? 2 : (c->mg_alloc ? 1 : 0);
    if (bytes) *bytes = c->mg_bytes;
    return MHDP_OK;
}

int mhdp_probe_next_cycle(mhdp_ctx* c, float smooth_weight, float coarse_scale, int coarse_visits) {
    if (!c || coarse_visits < 1 || coarse_visits > kMaxVisits || !std::isfinite(smooth_weight) ||
        !std::isfinite(coarse_scale))
        return MHDP_INVALID;
    c->test_cycle = Cycle{(double)smooth_weight, (double)coarse_scale, coarse_visits};
    return MHDP_OK;
}

int mhdp_probe_rhs(mhdp_ctx* c, int meanfree, const float* ustar, double* b) {
    if (!c || !ustar || !b) return MHDP_INVALID;
    int rc = prepare(c, meanfree != 0);
    if (rc) return rc;
    const size_t fb = face_count(c->grid) * sizeof(float);
    if ((rc = dalloc(c, (void**)&c->ustar_stage, fb)) || (rc = dalloc(c, (void**)&c->b64, (size_t)c->g.n * 8)))
        return rc;
    HIP_OK(hipMemcpyAsync(c->ustar_stage, ustar, fb, hipMemcpyHostToDevice, c->stream));
    launch_rhs_f64(c->stream, c->g, c->grid.h, c->mask, open8_of(c), c->bx, c->by, c->bz, c->ustar_stage, c->b64);
    HIP_OK(hipMemcpyAsync(b, c->b64, (size_t)c->g.n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return MHDP_OK;
}

}  // extern "C"
