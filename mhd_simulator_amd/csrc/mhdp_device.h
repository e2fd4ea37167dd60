// Device-side building blocks shared by the projection kernels (gfx950 / CDNA4).
//
// Layouts in HBM (one projection context):
//   cells      x + nx*(y + ny*z), 32-bit indices (grids up to 2^31 cells)
//   beta_a[i]  mobility of the LOWER face of cell i along axis a (between i-e_a, wrapped,
//              and i); 0 for a closed face.  The upper face of i is beta_a[i+e_a] (wrapped)
//              or 0 at a non-periodic boundary.
//   dinv[i]    1/diag for an active cell (a cell carrying pressure unknowns), 0 otherwise
//              (solid cells and, in the pinned gauge, each component's anchor cell).
//   faces      three SoA planes packed x|y|z, plane a has dims n + e_a.
//
// Work distribution: every cell-parallel kernel runs a fixed grid of blocks and gives each
// block one contiguous range of cells.  Hardware dispatches consecutive workgroups round-robin
// over the 8 XCDs, so the logical block index is remapped to keep each XCD on one contiguous
// z-slab of the grid (its y/z stencil neighbours then hit that XCD's own L2).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <stdlib.h>

#define MHDP_BLOCK 256
#define MHDP_XCDS 8

struct Grid3 {
    int nx, ny, nz;
    uint32_t per;  // periodic bits
    uint32_t n;    // cells
};

// Device scalar block — the reference's `scalars` buffer (solver.wgsl:11) in FP64,
// plus the solve parameters and the reduction tickets.
struct Scal {
    double rz, rr, bb, sr;            // scalars[0]: r.z, r.r, b.b, sum r
    double alpha, beta, conv, iters;  // scalars[1]
    double meanfree, mean;            // scalars[2].xy
    double tol2, abs2, bound, inv_nf; // Params.solve.xyz, Params.initial.w
    double pq, rr_stash, rz_old, spare;
    double omega, pscale;             // V-cycle smoother weight and coarse-correction scale
    double rr64, bb64;                // FP64 refinement: |b - L x|^2 and |b|^2
    double refine_tol, tau;           // FP64 target and the floor of the inner tolerance
    uint32_t anchor;                  // gauge cell of the single component (mean-free)
    uint32_t pending;                 // fused path: x += alpha p of the last iteration still owed
    uint32_t outer_done;              // FP64 refinement reached refine_tol (inner work is skipped)
    uint32_t rsel;                    // brick path: the residual lives in r (0) or its twin (1)
    double mean_next;                 // brick path: mean of the residual the next V-cycle sees
    uint32_t guard;                   // multigrid PCG: watch for a breakdown (r.z or p.Ap <= 0)
    uint32_t fault;                   // a watched breakdown happened (the solve then stops: conv = 1)
    uint32_t ticket[8];
};

enum Fin : int {
    FIN_NONE = 0,
    FIN_INITIAL,       // initial_reduce (solver.wgsl:198-203)
    FIN_REFRESH_MEAN,  // refresh partials + residual mean (mean-free refresh)
    FIN_INITIAL_SUB,   // initial_reduce after the initial mean-free subspace
    FIN_ALPHA,         // alpha_reduce (solver.wgsl:211-215)
    FIN_MEAN,          // pressure_residual_mean (pressure_subspace.wgsl:3-7)
    FIN_BETA,          // beta_reduce (solver.wgsl:231-239): rz=x, rr=y of this reduction
    FIN_STASH_RR,      // keep rr for the MG beta_reduce
    FIN_BETA_MG,       // beta_reduce with rz from preconditioned_dot and the stashed rr
    FIN_INITIAL_MG,    // initial_reduce after initial_direction
    FIN_CHECK,         // check_reduce (solver.wgsl:324-328)
    FIN_NORM64,        // outer refinement: rr of the FP64 residual
    FIN_UPDATE_MG,     // fused update: mean of r (mean-free) and |r - mean|^2 for beta_reduce
    FIN_OUTER,         // FP64 refinement step: outer test, inner tolerance and the inner start
    FIN_RESTRICT_UPD,  // restriction with the residual update: mean and |r - mean|^2 of the new r
};

__device__ __forceinline__ float ldf(float v) { return v; }
__device__ __forceinline__ float ldf(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ double ldd(double v) { return v; }
template <class T> __device__ __forceinline__ T cvt(float v);
template <> __device__ __forceinline__ float cvt<float>(float v) { return v; }
template <> __device__ __forceinline__ __hip_bfloat16 cvt<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

// Logical block (XCD-aware) and the cell range it owns.
__device__ __forceinline__ uint32_t logical_block() {
    uint32_t b = blockIdx.x, G = gridDim.x;
    uint32_t per = G / MHDP_XCDS;
    if (per * MHDP_XCDS != G) return b;
    return (b % MHDP_XCDS) * per + b / MHDP_XCDS;
}
__device__ __forceinline__ void block_range(uint32_t n, uint32_t& lo, uint32_t& hi) {
    uint32_t G = gridDim.x;
    uint32_t chunk = (n + G - 1) / G;
    chunk = (chunk + MHDP_BLOCK - 1) / MHDP_BLOCK * MHDP_BLOCK;
    uint32_t lb = logical_block();
    uint64_t l = (uint64_t)lb * chunk;
    uint64_t h = l + chunk;
    lo = l < n ? (uint32_t)l : n;
    hi = h < n ? (uint32_t)h : n;
}

// The blocks of one problem inside a launch: its logical block and the number of its blocks.
// A single-problem launch is (logical_block(), gridDim.x); a batched launch gives each problem
// a contiguous run of blocks and the same per-problem numbering, so results are identical.
struct VGrid {
    uint32_t b, n;
};
__device__ __forceinline__ VGrid vgrid() { return VGrid{logical_block(), gridDim.x}; }
__device__ __forceinline__ void block_range(const VGrid& vg, uint32_t n, uint32_t& lo, uint32_t& hi) {
    uint32_t chunk = (n + vg.n - 1) / vg.n;
    chunk = (chunk + MHDP_BLOCK - 1) / MHDP_BLOCK * MHDP_BLOCK;
    uint64_t l = (uint64_t)vg.b * chunk;
    uint64_t h = l + chunk;
    lo = l < n ? (uint32_t)l : n;
    hi = h < n ? (uint32_t)h : n;
}

__device__ __forceinline__ void decode(const Grid3& g, uint32_t i, int& x, int& y, int& z) {
    uint32_t nxy = (uint32_t)g.nx * (uint32_t)g.ny;
    z = (int)(i / nxy);
    uint32_t r = i - (uint32_t)z * nxy;
    y = (int)(r / (uint32_t)g.nx);
    x = (int)(r - (uint32_t)y * (uint32_t)g.nx);
}

// Neighbour indices and face weights of cell i in the reference's summation order
// x-, x+, y-, y+, z-, z+ (solver.wgsl:81).  Indices are always in range: a closed
// boundary face gets weight 0 and the cell itself as neighbour.
template <class CT>
__device__ __forceinline__ void stencil6(const Grid3& g, int x, int y, int z, uint32_t i,
                                         const CT* __restrict__ bx, const CT* __restrict__ by,
                                         const CT* __restrict__ bz, uint32_t nb[6], float w[6]) {
    const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
    // x
    nb[0] = x > 0 ? i - 1 : ((g.per & 1u) ? i + (uint32_t)(g.nx - 1) : i);
    w[0] = ldf(bx[i]);
    if (x + 1 < g.nx) { nb[1] = i + 1; w[1] = ldf(bx[i + 1]); }
    else if (g.per & 1u) { nb[1] = i - (uint32_t)(g.nx - 1); w[1] = ldf(bx[nb[1]]); }
    else { nb[1] = i; w[1] = 0.f; }
    // y
    nb[2] = y > 0 ? i - sy : ((g.per & 2u) ? i + (uint32_t)(g.ny - 1) * sy : i);
    w[2] = ldf(by[i]);
    if (y + 1 < g.ny) { nb[3] = i + sy; w[3] = ldf(by[i + sy]); }
    else if (g.per & 2u) { nb[3] = i - (uint32_t)(g.ny - 1) * sy; w[3] = ldf(by[nb[3]]); }
    else { nb[3] = i; w[3] = 0.f; }
    // z
    nb[4] = z > 0 ? i - sz : ((g.per & 4u) ? i + (uint32_t)(g.nz - 1) * sz : i);
    w[4] = ldf(bz[i]);
    if (z + 1 < g.nz) { nb[5] = i + sz; w[5] = ldf(bz[i + sz]); }
    else if (g.per & 4u) { nb[5] = i - (uint32_t)(g.nz - 1) * sz; w[5] = ldf(bz[nb[5]]); }
    else { nb[5] = i; w[5] = 0.f; }
}

// ---------------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Fixed-order block sum of NV lanes over a block of NT threads; thread 0 receives the totals.
template <int NV, int NT = MHDP_BLOCK>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[NT / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
            for (int w = 0; w < NT / 64; w++) s += red[w][k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// The scalar updates of the reference's single-workgroup reduction kernels.
__device__ inline void finalize(Scal* s, int mode, const double t[4]) {
    switch (mode) {
    case FIN_INITIAL:  // initial_reduce: scalars[0] = total; scalars[1] = (0, 0, conv, iters)
        s->rz = t[0]; s->rr = t[1]; s->bb = t[2]; s->sr = t[3];
        s->alpha = 0.0; s->beta = 0.0;
        s->mean = 0.0;
        s->pending = 0u;
        s->rsel = 0u;
        s->conv = (t[1] <= s->tol2 * t[2] + s->abs2 || t[2] == 0.0) ? 1.0 : 0.0;
        break;
    case FIN_REFRESH_MEAN:  // keep the physical residual and b.b; mean of r for the subspace step
        s->rr = t[1]; s->bb = t[2]; s->sr = t[3];
        s->mean = t[3] * s->inv_nf;
        s->rsel = 0u;
        break;
    case FIN_INITIAL_SUB:  // initial_reduce after pressure_initial_subspace (old.y, old.z kept)
        s->rz = t[0];
        s->alpha = 0.0; s->beta = 0.0;
        s->mean = 0.0;  // r is projected now
        s->sr = 0.0;
        s->rsel = 0u;
        s->pending = 0u;
        s->conv = (s->rr <= s->tol2 * s->bb + s->abs2 || s->bb == 0.0) ? 1.0 : 0.0;
        break;
    case FIN_ALPHA:  // t[1] = sum of q over the active cells (brick apply)
        s->pq = t[0];
        if (s->guard && s->conv < 0.5 && t[0] <= 0.0) {  // p.Ap <= 0: the preconditioner is not SPD
            s->fault = 1u;
            s->conv = 1.0;
        }
        s->alpha = s->conv > 0.5 ? 0.0 : s->rz / fmax(t[0], 1e-35);
        s->pending = 1u;
        // mean of the next residual r' = (r - mean) - alpha q, known before r' exists
        s->mean_next = (s->sr * s->inv_nf - s->mean) - s->alpha * t[1] * s->inv_nf;
        break;
    case FIN_UPDATE_MG:  // sum r^2 and sum r of the updated (unprojected) residual
        s->mean = t[3] * s->inv_nf;
        s->sr = t[3];
        s->rr_stash = t[1] - s->mean * t[3];
        break;
    case FIN_RESTRICT_UPD:  // as FIN_UPDATE_MG; the new residual is in the other buffer
        s->mean = t[3] * s->inv_nf;
        s->sr = t[3];
        s->rr_stash = t[1] - s->mean * t[3];
        s->rsel ^= 1u;
        break;
    case FIN_MEAN:
        s->mean = t[3] * s->inv_nf;
        s->rr_stash = t[1];
        break;
    case FIN_BETA:
    case FIN_BETA_MG:
        if (s->conv < 0.5) {
            double rr = mode == FIN_BETA ? t[1] : s->rr_stash;
            s->iters += 1.0;
            s->rz_old = s->rz;
            s->beta = t[0] / fmax(s->rz, 1e-35);
            s->rz = t[0];
            s->rr = rr;
            s->conv = rr <= (s->tol2 * s->bb + s->abs2) / s->bound ? 1.0 : 0.0;
            // r.z <= 0, or a residual grown 1e5-fold over the solve's right-hand side: the cycle is
            // not positive definite on this problem
            if (mode == FIN_BETA_MG && s->guard && s->conv < 0.5 && (t[0] <= 0.0 || rr > 1e10 * s->bb)) {
                s->fault = 1u;
                s->conv = 1.0;
            }
        }
        break;
    case FIN_STASH_RR:
        s->rr_stash = t[1];
        break;
    case FIN_INITIAL_MG:  // initial_reduce after initial_direction: x lane = r.z
        s->rz = t[0];
        s->alpha = 0.0; s->beta = 0.0;
        s->pending = 0u;
        s->conv = (s->rr <= s->tol2 * s->bb + s->abs2 || s->bb == 0.0) ? 1.0 : 0.0;
        if (s->guard && s->conv < 0.5 && t[0] <= 0.0) {  // r.z <= 0 for r != 0
            s->fault = 1u;
            s->conv = 1.0;
        }
        break;
    case FIN_CHECK:
        s->rr = t[1];
        s->conv = t[1] <= s->tol2 * s->bb + s->abs2 ? 1.0 : 0.0;
        break;
    case FIN_NORM64:  // FP64 outer residual: rr, b.b (in `spare`), sum r
        s->rr = t[1];
        s->spare = t[2];
        s->sr = t[3];
        break;
    case FIN_OUTER: {  // t = (sum r32^2, |b - L x|^2, |b|^2, sum r32) of the FP64 residual r
        s->rr64 = t[1];
        s->bb64 = t[2];
        const double res = t[2] > 0.0 ? sqrt(t[1] / t[2]) : 0.0;
        const bool done = t[2] == 0.0 || res <= s->refine_tol;
        // inner tolerance: no tighter than the remaining factor asks for, at most 0.5
        double tl = res > 0.0 ? fmax(s->tau, 0.5 * s->refine_tol / res) : 0.5;
        if (tl > 0.5) tl = 0.5;
        s->tol2 = tl * tl;
        s->abs2 = 0.0;
        s->bound = 1.0;
        // the inner solve starts from x = 0 with b = r32; its mean-free projection is
        // deferred (the next passes read r - mean)
        const double m = s->meanfree > 0.5 ? t[3] * s->inv_nf : 0.0;
        s->mean = m;
        s->mean_next = m;
        s->rsel = 0u;
        s->bb = t[0];
        s->rr = t[0] - m * t[3];
        s->sr = t[3];
        s->rz = 0.0; s->alpha = 0.0; s->beta = 0.0; s->iters = 0.0;
        s->pending = 0u;
        s->outer_done = done ? 1u : 0u;
        s->conv = (done || s->rr <= s->tol2 * s->bb || s->bb == 0.0) ? 1.0 : 0.0;
        break;
    }
    default:
        break;
    }
}

// Every block stores its partial; the last block to finish folds all partials in a
// fixed order (deterministic) and runs the scalar update `mode` — one launch per
// reference reduction pair (e.g. pressure_apply + alpha_reduce).
// The partials and the ticket are device-scope atomic accesses (they bypass the per-XCD L2),
// and a block waits for its partial stores to complete before taking its ticket.  No
// __threadfence(): on this device a device-scope fence writes back AND invalidates the XCD's
// L2, which, once per block, throws away the halo data the other blocks are reading.
template <int NV, int NT = MHDP_BLOCK>
__device__ __forceinline__ void reduce_finish(const VGrid& vg, double (&v)[NV], double* __restrict__ part, Scal* s,
                                              int ticket, int mode) {
    block_sum<NV, NT>(v);
    __shared__ int last;
    const uint32_t lb = vg.b;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            __hip_atomic_store(part + 4 * lb + k, k < NV ? v[k] : 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);  // the partial stores have completed
        const unsigned t = __hip_atomic_fetch_add(&s->ticket[ticket], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = (t == vg.n - 1);
    }
    __syncthreads();
    if (!last) return;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t j = threadIdx.x; j < vg.n; j += NT) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            acc[k] += __hip_atomic_load(part + 4 * j + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<4, NT>(acc);
    if (threadIdx.x == 0) {
        finalize(s, mode, acc);
        s->ticket[ticket] = 0u;
    }
}

template <int NV>
__device__ __forceinline__ void reduce_finish(double (&v)[NV], double* __restrict__ part, Scal* s, int ticket,
                                              int mode) {
    reduce_finish<NV>(vgrid(), v, part, s, ticket, mode);
}

// ---------------------------------------------------------------- launch helpers (host)

struct Launch {
    uint32_t blocks;
};
// Launch shapes of the cell-parallel and stencil kernels, process-wide: read from the
// environment at first use (MHDP_GRID_CAP, MHDP_BRICK_CAP, MHDP_MARCH_CAP, MHDP_MARCH_Z,
// MHDP_BRICK_OCC) and replaced by mhdp_set_launch_tuning (the autotuner).
struct LaunchTuning {
    uint32_t grid_cap;   // block cap of a cell-parallel launch (each block owns a contiguous range)
    uint32_t brick_cap;  // block cap of the 4x2x2 brick stencil kernels
    uint32_t march_cap;  // block cap of the z-marching kernels
    uint32_t march_z;    // planes per z-marching work item (even)
    int brick_occ;       // resident blocks per CU the brick kernels' registers target (2, 3, 4)
};
LaunchTuning& launch_tuning();
static inline uint32_t grid_cap() { return launch_tuning().grid_cap; }
static inline uint32_t brick_grid_cap() { return launch_tuning().brick_cap; }
static inline uint32_t grid_blocks(uint32_t n, uint32_t cap = 0) {
    if (cap == 0) cap = grid_cap();
    uint32_t need = (n + MHDP_BLOCK - 1) / MHDP_BLOCK;
    uint32_t b = need < cap ? need : cap;
    if (b >= 64) b = (b + MHDP_XCDS - 1) / MHDP_XCDS * MHDP_XCDS;
    return b ? b : 1;
}
