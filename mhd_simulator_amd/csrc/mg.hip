// Aggregation multigrid V(1,1) preconditioner (multigrid.wgsl): weighted-Jacobi (omega = 2/3)
// smoothing, restriction by summing the 2x2x2 children's residuals, piecewise-constant
// prolongation and a fixed 8-sweep Jacobi solve on the coarsest grid (<= 4^3 cells, one
// workgroup).  Optionally level 1 is visited more than once per cycle (a W-cycle there:
// coarse_visits_t).  Templated on the storage type of coefficients (CT) and vectors (VT):
// float for the FP32 solver, bfloat16 for the mixed-precision solver; arithmetic is FP32
// (FP64 on the coarsest grid in the mixed solver).
#include "mhdp_kernels.h"
#include "launch.h"

#define CONVERGED(sc) ((sc)->conv > 0.5)

namespace {

__device__ __forceinline__ int wrapc(int v, int n, bool periodic) {
    if (!periodic) return v;
    return v < 0 ? v + n : (v >= n ? v - n : v);
}

// Coarse-graph action at cell i: react*z_i + sum_valid beta (z_i - z_n), with beta the
// lower-slot edge of i (s<0) or of the neighbour (s>0) (action(), multigrid.wgsl:44-48).
template <class CT, class VT>
__device__ __forceinline__ float coarse_action(const Grid3& g, int x, int y, int z, uint32_t i,
                                               const CT* __restrict__ cb0, const CT* __restrict__ cb1,
                                               const CT* __restrict__ cb2, const CT* __restrict__ react,
                                               const VT* __restrict__ v, float center) {
    const int d[3] = {g.nx, g.ny, g.nz};
    const CT* B[3] = {cb0, cb1, cb2};
    int q[3] = {x, y, z};
    float out = ldf(react[i]) * center;
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int s = -1; s <= 1; s += 2) {
            int n[3] = {q[0], q[1], q[2]};
            n[a] = wrapc(n[a] + s, d[a], (g.per >> a) & 1u);
            if (n[a] < 0 || n[a] >= d[a]) continue;
            const uint32_t ni = (uint32_t)n[0] + (uint32_t)g.nx * ((uint32_t)n[1] + (uint32_t)g.ny * (uint32_t)n[2]);
            const float beta = ldf(s > 0 ? B[a][ni] : B[a][i]);
            out += beta * (center - ldf(v[ni]));
        }
    }
    return out;
}

template <class CT, class VT>
__device__ __forceinline__ float fine_action(const Grid3& g, int x, int y, int z, uint32_t i,
                                             const CT* __restrict__ bx, const CT* __restrict__ by,
                                             const CT* __restrict__ bz, const VT* __restrict__ v, float center) {
    uint32_t nb[6];
    float w[6];
    stencil6(g, x, y, z, i, bx, by, bz, nb, w);
    float out = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; k++) out += w[k] * (center - ldf(v[nb[k]]));
    return out;
}

}  // namespace

// diagonal_start (multigrid.wgsl:134-139): z = omega r / d from a zero guess.
template <class CT, class RT, class VT>
struct K_k_dstart {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const RT* r;
        const CT* dinv;
        VT* z;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        uint32_t n = A_.n;
        const RT* __restrict__ r = A_.r;
        const CT* __restrict__ dinv = A_.dinv;
        VT* __restrict__ z = A_.z;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
        uint32_t lo, hi;
        block_range(vg, n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const float di = ldf(dinv[i]);
            z[i] = cvt<VT>(di != 0.0f ? omega * ldf(r[i]) * di : 0.0f);
        }
    }
};

// restrict_residual (multigrid.wgsl:75-83) from the fine level: sum over active children
// of r - A z.
template <class CT, class VT>
struct K_k_restrict_fine {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 cg;
        Grid3 fg;
        const float* r;
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
        const VT* z;
        VT* rc;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 cg = A_.cg;
        Grid3 fg = A_.fg;
        const float* __restrict__ r = A_.r;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ bx = A_.bx;
        const CT* __restrict__ by = A_.by;
        const CT* __restrict__ bz = A_.bz;
        const VT* __restrict__ z = A_.z;
        VT* __restrict__ rc = A_.rc;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, cg.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int qx, qy, qz;
            decode(cg, i, qx, qy, qz);
            float rhs = 0.0f;
            for (int dz = 0; dz < 2; dz++)
                for (int dy = 0; dy < 2; dy++)
                    for (int dx = 0; dx < 2; dx++) {
                        const int x = 2 * qx + dx, y = 2 * qy + dy, zz = 2 * qz + dz;
                        if (x >= fg.nx || y >= fg.ny || zz >= fg.nz) continue;
                        const uint32_t ci = (uint32_t)x + (uint32_t)fg.nx * ((uint32_t)y + (uint32_t)fg.ny * (uint32_t)zz);
                        if (ldf(dinv[ci]) == 0.0f) continue;
                        rhs += r[ci] - fine_action(fg, x, y, zz, ci, bx, by, bz, z, ldf(z[ci]));
                    }
            rc[i] = cvt<VT>(rhs);
        }
    }
};

// restrict_residual between coarse levels: every child contributes.
template <class CT, class VT>
struct K_k_restrict_coarse {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 cg;
        Grid3 fg;
        const VT* rf;
        const CT* cb0;
        const CT* cb1;
        const CT* cb2;
        const CT* react;
        const VT* zf;
        VT* rc;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 cg = A_.cg;
        Grid3 fg = A_.fg;
        const VT* __restrict__ rf = A_.rf;
        const CT* __restrict__ cb0 = A_.cb0;
        const CT* __restrict__ cb1 = A_.cb1;
        const CT* __restrict__ cb2 = A_.cb2;
        const CT* __restrict__ react = A_.react;
        const VT* __restrict__ zf = A_.zf;
        VT* __restrict__ rc = A_.rc;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, cg.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int qx, qy, qz;
            decode(cg, i, qx, qy, qz);
            float rhs = 0.0f;
            for (int dz = 0; dz < 2; dz++)
                for (int dy = 0; dy < 2; dy++)
                    for (int dx = 0; dx < 2; dx++) {
                        const int x = 2 * qx + dx, y = 2 * qy + dy, zz = 2 * qz + dz;
                        if (x >= fg.nx || y >= fg.ny || zz >= fg.nz) continue;
                        const uint32_t ci = (uint32_t)x + (uint32_t)fg.nx * ((uint32_t)y + (uint32_t)fg.ny * (uint32_t)zz);
                        rhs += ldf(rf[ci]) - coarse_action(fg, x, y, zz, ci, cb0, cb1, cb2, react, zf, ldf(zf[ci]));
                    }
            rc[i] = cvt<VT>(rhs);
        }
    }
};

// prolong (multigrid.wgsl:84-88): z_child += z_parent (fine children: active only).
template <class CT, class VT>
struct K_k_prolong {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 fg;
        Grid3 cg;
        int fine;
        const VT* zc;
        const CT* dinv;
        VT* zf;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 fg = A_.fg;
        Grid3 cg = A_.cg;
        int fine = A_.fine;
        const VT* __restrict__ zc = A_.zc;
        const CT* __restrict__ dinv = A_.dinv;
        VT* __restrict__ zf = A_.zf;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, fg.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int x, y, z;
            decode(fg, i, x, y, z);
            if (fine && ldf(dinv[i]) == 0.0f) continue;
            const uint32_t pi =
                (uint32_t)(x / 2) + (uint32_t)cg.nx * ((uint32_t)(y / 2) + (uint32_t)cg.ny * (uint32_t)(z / 2));
            zf[i] = cvt<VT>(ldf(zf[i]) + (float)sc->pscale * ldf(zc[pi]));
        }
    }
};

// smooth_to_temp (multigrid.wgsl:66-73): t = z + omega (r - A z) / d, 0 where inactive.
template <class CT, class RT, class VT, bool FINE>
struct K_k_smooth {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const RT* r;
        const CT* dinv;
        const CT* b0;
        const CT* b1;
        const CT* b2;
        const CT* react;
        const VT* z;
        VT* t;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const RT* __restrict__ r = A_.r;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ b0 = A_.b0;
        const CT* __restrict__ b1 = A_.b1;
        const CT* __restrict__ b2 = A_.b2;
        const CT* __restrict__ react = A_.react;
        const VT* __restrict__ z = A_.z;
        VT* __restrict__ t = A_.t;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const float di = ldf(dinv[i]);
            float out = 0.0f;
            if (di != 0.0f) {
                int x, y, zz;
                decode(g, i, x, y, zz);
                const float c = ldf(z[i]);
                const float a = FINE ? fine_action(g, x, y, zz, i, b0, b1, b2, z, c)
                                     : coarse_action(g, x, y, zz, i, b0, b1, b2, react, z, c);
                out = c + omega * (ldf(r[i]) - a) * di;
            }
            t[i] = cvt<VT>(out);
        }
    }
};

// coarse_solve (multigrid.wgsl:104-131): COARSE_SWEEPS = 8 weighted-Jacobi sweeps from zero
// on the coarsest level (<= 256 cells) inside one workgroup.  AT is the arithmetic type.
template <class CT, class RT, class VT, class AT, bool FINE>
struct K_k_coarse {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const RT* r;
        const CT* dinv;
        const CT* b0;
        const CT* b1;
        const CT* b2;
        const CT* react;
        VT* zout;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const RT* __restrict__ r = A_.r;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ b0 = A_.b0;
        const CT* __restrict__ b1 = A_.b1;
        const CT* __restrict__ b2 = A_.b2;
        const CT* __restrict__ react = A_.react;
        VT* __restrict__ zout = A_.zout;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        __shared__ AT cs[MHDP_BLOCK];
        const uint32_t i = threadIdx.x;
        const bool in_grid = i < g.n;
        AT rhs = 0, di = 0, rc = 0;
        int q[3] = {0, 0, 0};
        if (in_grid) {
            decode(g, i, q[0], q[1], q[2]);
            di = (AT)ldf(dinv[i]);
            rhs = di != (AT)0 ? (AT)ldf(r[i]) : (AT)0;
            if (!FINE) rc = (AT)ldf(react[i]);
        }
        cs[i] = 0;
        __syncthreads();
        const int d[3] = {g.nx, g.ny, g.nz};
        const CT* B[3] = {b0, b1, b2};
        for (int step = 0; step < 8; step++) {
            const AT center = cs[i];
            AT next = 0;
            if (in_grid && di != (AT)0) {
                AT a = rc * center;
                for (int ax = 0; ax < 3; ax++)
                    for (int s = -1; s <= 1; s += 2) {
                        int n[3] = {q[0], q[1], q[2]};
                        n[ax] = wrapc(n[ax] + s, d[ax], (g.per >> ax) & 1u);
                        if (n[ax] < 0 || n[ax] >= d[ax]) continue;
                        const uint32_t ni =
                            (uint32_t)n[0] + (uint32_t)g.nx * ((uint32_t)n[1] + (uint32_t)g.ny * (uint32_t)n[2]);
                        AT nbv = cs[ni];
                        if (FINE && ldf(dinv[ni]) == 0.0f) nbv = 0;
                        const AT beta = (AT)ldf(s > 0 ? B[ax][ni] : B[ax][i]);
                        a += beta * (center - nbv);
                    }
                next = center + (AT)sc->omega * (rhs - a) * di;
            }
            __syncthreads();
            cs[i] = next;
            __syncthreads();
        }
        if (in_grid) zout[i] = cvt<VT>((float)cs[i]);
    }
};

template <class CT, class VT, class AT>
static void coarse_visits_t(hipStream_t s, int L, LevelDev* lv, Scal* sc, int visits, void** z1out, bool f16,
                            bool rb);

template <class CT, class VT, class AT>
static void vcycle_t(hipStream_t s, int L, LevelDev* lv, const float* r0, Scal* sc, void** zfinal, bool f16,
                     int visits, KernelTimer* timer) {
    KernelTimer dummy;
    KernelTimer& T = timer ? *timer : dummy;
    auto C = [&](LevelDev& l, int which) -> const CT* {
        if (f16) {
            switch (which) {
            case 0: return (const CT*)l.bx16;
            case 1: return (const CT*)l.by16;
            case 2: return (const CT*)l.bz16;
            case 3: return (const CT*)l.react16;
            default: return (const CT*)l.dinv16;
            }
        }
        switch (which) {
        case 0: return (const CT*)l.bx;
        case 1: return (const CT*)l.by;
        case 2: return (const CT*)l.bz;
        case 3: return (const CT*)l.react;
        default: return (const CT*)l.dinv;
        }
    };
    const dim3 B(MHDP_BLOCK);
    if (L == 1) {
        LevelDev& f = lv[0];
        go<K_k_coarse<CT, float, VT, AT, true>>(s, 1, {f.g, r0, C(f, 4), C(f, 0), C(f, 1), C(f, 2), (const CT*)nullptr, (VT*)f.z, sc});
        *zfinal = f.z;
        return;
    }
    if (visits >= 2 && L >= 3) {  // the fine level around several visits of level 1
        LevelDev& a = lv[0];
        LevelDev& c = lv[1];
        go<K_k_dstart<CT, float, VT>>(s, grid_blocks(a.g.n), {a.g.n, r0, C(a, 4), (VT*)a.z, sc});
        T.before(3, s);
        go<K_k_restrict_fine<CT, VT>>(s, grid_blocks(c.g.n), {c.g, a.g, r0, C(a, 4), C(a, 0), C(a, 1), C(a, 2), (const VT*)a.z, (VT*)c.r, sc});
        T.after(3, s);
        void* z1 = nullptr;
        coarse_visits_t<CT, VT, AT>(s, L, lv, sc, visits, &z1, f16, false);
        go<K_k_prolong<CT, VT>>(s, grid_blocks(a.g.n), {a.g, c.g, 1, (const VT*)z1, C(a, 4), (VT*)a.z, sc});
        T.before(2, s);
        go<K_k_smooth<CT, float, VT, true>>(s, grid_blocks(a.g.n), {a.g, r0, C(a, 4), C(a, 0), C(a, 1), C(a, 2), (const CT*)nullptr, (const VT*)a.z, (VT*)a.t, sc});
        T.after(2, s);
        *zfinal = a.t;
        return;
    }
    // descend
    for (int l = 0; l < L - 1; l++) {
        LevelDev& a = lv[l];
        LevelDev& c = lv[l + 1];
        if (l == 0) {
            go<K_k_dstart<CT, float, VT>>(s, grid_blocks(a.g.n), {a.g.n, r0, C(a, 4), (VT*)a.z, sc});
            T.before(3, s);
            go<K_k_restrict_fine<CT, VT>>(s, grid_blocks(c.g.n), {c.g, a.g, r0, C(a, 4), C(a, 0), C(a, 1), C(a, 2), (const VT*)a.z, (VT*)c.r, sc});
            T.after(3, s);
        } else {
            go<K_k_dstart<CT, VT, VT>>(s, grid_blocks(a.g.n), {a.g.n, (const VT*)a.r, C(a, 4), (VT*)a.z, sc});
            go<K_k_restrict_coarse<CT, VT>>(s, grid_blocks(c.g.n), {c.g, a.g, (const VT*)a.r, C(a, 0), C(a, 1), C(a, 2), C(a, 3), (const VT*)a.z, (VT*)c.r, sc});
        }
    }
    {
        LevelDev& c = lv[L - 1];
        go<K_k_coarse<CT, VT, VT, AT, false>>(s, 1, {c.g, (const VT*)c.r, C(c, 4), C(c, 0), C(c, 1), C(c, 2), C(c, 3), (VT*)c.z, sc});
    }
    // ascend
    for (int l = L - 2; l >= 0; l--) {
        LevelDev& a = lv[l];
        LevelDev& c = lv[l + 1];
        const VT* zc = (const VT*)(l + 1 == L - 1 ? c.z : c.t);
        go<K_k_prolong<CT, VT>>(s, grid_blocks(a.g.n), {a.g, c.g, l == 0 ? 1 : 0, zc, C(a, 4), (VT*)a.z, sc});
        if (l == 0) {
            T.before(2, s);
            go<K_k_smooth<CT, float, VT, true>>(s, grid_blocks(a.g.n), {a.g, r0, C(a, 4), C(a, 0), C(a, 1), C(a, 2), (const CT*)nullptr, (const VT*)a.z, (VT*)a.t, sc});
            T.after(2, s);
        } else
            go<K_k_smooth<CT, VT, VT, false>>(s, grid_blocks(a.g.n), {a.g, (const VT*)a.r, C(a, 4), C(a, 0), C(a, 1), C(a, 2), C(a, 3), (const VT*)a.z, (VT*)a.t, sc});
    }
    *zfinal = lv[0].t;
}

// r = r - A z on a coarse level, in place (each cell reads only its own r): the residual that one
// visit of the level leaves, the right-hand side of the next visit (W-cycle).
template <class CT, class VT>
struct K_k_coarse_residual {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const CT *bx, *by, *bz, *react;
        const VT* z;
        VT* r;
        Scal* sc;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (CONVERGED(a.sc)) return;
        const Grid3 g = a.g;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        const uint32_t nxy = (uint32_t)g.nx * (uint32_t)g.ny;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const int z = (int)(i / nxy), rem = (int)(i - (uint32_t)z * nxy);
            const int y = rem / g.nx, x = rem - y * g.nx;
            const float act = coarse_action(g, x, y, z, i, a.bx, a.by, a.bz, a.react, a.z, ldf(a.z[i]));
            a.r[i] = cvt<VT>(ldf(a.r[i]) - act);
        }
    }
};
// t = t + theta w: the corrections of two visits of a level, the later one damped
template <class VT>
struct K_k_add {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        float theta;
        const VT* w;
        VT* t;
        Scal* sc;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (CONVERGED(a.sc)) return;
        uint32_t lo, hi;
        block_range(vg, a.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK)
            a.t[i] = cvt<VT>(ldf(a.t[i]) + a.theta * ldf(a.w[i]));
    }
};

// Levels 1..L-1 of the V-cycle (L >= 2) from the level-1 residual in lv[1].r.  acc (brick level 1
// and L >= 3 only): the level-1 smoother writes acc + theta * (its result) into acc instead of
// lv[1].t, and acc is the output.
// How a coarse level (not the coarsest) is smoothed: red-black passes, 4x2x2 Jacobi bricks, or the
// one-cell Jacobi kernels.  The first two can add the level's result into an accumulator.
enum LevelKind { LK_RB, LK_BRICK, LK_CELL };
static LevelKind level_kind(const LevelDev& l, bool rb) {
    if (rb && rb_level_ok(l.g)) return LK_RB;
    if (brick_enabled_env() && brick_ok(l.g)) return LK_BRICK;
    return LK_CELL;
}

template <class CT, class VT, class AT>
static void vcycle_coarse_t(hipStream_t s, int L, LevelDev* lv, Scal* sc, void** z1out, bool f16, bool rb,
                            void* acc = nullptr, float theta = 0.0f) {
    auto C = [&](LevelDev& l, int which) -> const CT* {
        void* const a16[5] = {l.bx16, l.by16, l.bz16, l.react16, l.dinv16};
        float* const a32[5] = {l.bx, l.by, l.bz, l.react, l.dinv};
        return f16 ? (const CT*)a16[which] : (const CT*)a32[which];
    };
    const dim3 B(MHDP_BLOCK);
    const int store = f16 ? STORE_BF16 : STORE_F32;
    // the levels from T down, with the coarsest solve, run as one launch
    const int T = rb ? rb_tail_level(L, lv) : L;
    const int down = T < L ? T : L - 1;
    for (int l = 1; l < down; l++) {
        LevelDev& a = lv[l];
        LevelDev& c = lv[l + 1];
        const LevelKind kind = level_kind(a, rb);
        if (kind == LK_RB) {
            launch_rb_restrict(s, store, true, c.g, a.g, a.r, a.r, nullptr, nullptr, nullptr, C(a, 4), C(a, 0),
                               C(a, 1), C(a, 2), c.r, nullptr, sc);
            continue;
        }
        if (kind == LK_BRICK) {
            launch_brick_restrict_coarse(s, store, c.g, a.g, a.r, C(a, 4), C(a, 0), C(a, 1), C(a, 2), C(a, 3), c.r, sc);
            continue;
        }
        go<K_k_dstart<CT, VT, VT>>(s, grid_blocks(a.g.n), {a.g.n, (const VT*)a.r, C(a, 4), (VT*)a.z, sc});
        go<K_k_restrict_coarse<CT, VT>>(s, grid_blocks(c.g.n), {c.g, a.g, (const VT*)a.r, C(a, 0), C(a, 1), C(a, 2), C(a, 3), (const VT*)a.z, (VT*)c.r, sc});
    }
    if (T < L) {
        launch_rb_tail(s, store, sizeof(AT) == 8, L, lv, T, sc);
    } else {
        LevelDev& c = lv[L - 1];
        go<K_k_coarse<CT, VT, VT, AT, false>>(s, 1, {c.g, (const VT*)c.r, C(c, 4), C(c, 0), C(c, 1), C(c, 2), C(c, 3), (VT*)c.z, sc});
    }
    for (int l = down - 1; l >= 1; l--) {
        LevelDev& a = lv[l];
        LevelDev& c = lv[l + 1];
        const VT* zc = (const VT*)(l + 1 == L - 1 ? c.z : c.t);
        const LevelKind kind = level_kind(a, rb);
        if (kind == LK_RB) {
            void* const into = (l == 1 && acc) ? acc : a.t;
            launch_rb_smooth(s, store, true, a.g, c.g, a.r, a.r, C(a, 4), C(a, 0), C(a, 1), C(a, 2), zc, into,
                             l == 1 ? acc : nullptr, theta, nullptr, sc, 0, false);
            continue;
        }
        if (kind == LK_BRICK) {
            void* const into = (l == 1 && acc) ? acc : a.t;
            launch_brick_smooth_coarse(s, store, a.g, c.g, a.r, C(a, 4), C(a, 0), C(a, 1), C(a, 2), C(a, 3), zc, into,
                                       sc, l == 1 ? acc : nullptr, theta);
            continue;
        }
        go<K_k_prolong<CT, VT>>(s, grid_blocks(a.g.n), {a.g, c.g, 0, zc, C(a, 4), (VT*)a.z, sc});
        go<K_k_smooth<CT, VT, VT, false>>(s, grid_blocks(a.g.n), {a.g, (const VT*)a.r, C(a, 4), C(a, 0), C(a, 1), C(a, 2), C(a, 3), (const VT*)a.z, (VT*)a.t, sc});
    }
    *z1out = (L - 1 == 1) ? lv[1].z : (acc ? acc : lv[1].t);
}

// The coarse part of the cycle, visiting level 1 `visits` times: after each visit the level-1
// right-hand side becomes the residual that visit leaves (in place) and the next visit's
// correction, computed into lv[1].w, is added to the first's, damped by kSecondVisit.  With M the
// level-1 cycle and A its operator, two visits apply (1 + theta) M - theta M A M, which is positive
// definite while the eigenvalues of M A stay below (1 + theta) / theta: 2.25 at theta = 0.8, where
// the over-corrected cycle (coarse scale 1.6) reaches about 1.75 on the droplet and plate grids.
constexpr float kSecondVisit = 0.8f;
template <class CT, class VT, class AT>
static void coarse_visits_t(hipStream_t s, int L, LevelDev* lv, Scal* sc, int visits, void** z1out, bool f16,
                            bool rb) {
    vcycle_coarse_t<CT, VT, AT>(s, L, lv, sc, z1out, f16, rb);
    if (visits < 2 || L < 3 || !lv[1].w) return;
    LevelDev& a = lv[1];
    const CT* B[4] = {(const CT*)(f16 ? a.bx16 : (void*)a.bx), (const CT*)(f16 ? a.by16 : (void*)a.by),
                      (const CT*)(f16 ? a.bz16 : (void*)a.bz), (const CT*)(f16 ? a.react16 : (void*)a.react)};
    const int store = f16 ? STORE_BF16 : STORE_F32;
    const bool brick = brick_enabled_env() && brick_ok(a.g);
    const bool sums = level_kind(a, rb) != LK_CELL;  // the level-1 smoother adds into the sum itself
    void* const sum = a.t;  // L >= 3: the first visit's correction is in lv[1].t
    const void* last = sum;
    for (int v = 1; v < visits; v++) {
        if (brick)
            launch_brick_coarse_residual(s, store, a.g, B[0], B[1], B[2], B[3], last, a.r, sc);
        else
            go<K_k_coarse_residual<CT, VT>>(s, grid_blocks(a.g.n), {a.g, B[0], B[1], B[2], B[3], (const VT*)last, (VT*)a.r, sc});
        void* zv = nullptr;
        if (sums && v == visits - 1) {  // the last visit's level-1 smoother adds into the sum itself
            vcycle_coarse_t<CT, VT, AT>(s, L, lv, sc, &zv, f16, rb, sum, kSecondVisit);
            break;
        }
        a.t = a.w;
        vcycle_coarse_t<CT, VT, AT>(s, L, lv, sc, &zv, f16, rb);
        a.t = sum;
        go<K_k_add<VT>>(s, grid_blocks(a.g.n), {a.g.n, kSecondVisit, (const VT*)zv, (VT*)sum, sc});
        last = zv;
    }
    *z1out = sum;
}

void launch_vcycle_coarse(hipStream_t s, int store, int L, LevelDev* lv, Scal* sc, int coarse_f64, int visits,
                          void** z1out, bool rb) {
    if (store == STORE_BF16) {
        if (coarse_f64) coarse_visits_t<__hip_bfloat16, __hip_bfloat16, double>(s, L, lv, sc, visits, z1out, true, rb);
        else coarse_visits_t<__hip_bfloat16, __hip_bfloat16, float>(s, L, lv, sc, visits, z1out, true, rb);
    } else {
        if (coarse_f64) coarse_visits_t<float, float, double>(s, L, lv, sc, visits, z1out, false, rb);
        else coarse_visits_t<float, float, float>(s, L, lv, sc, visits, z1out, false, rb);
    }
}

void launch_vcycle(hipStream_t s, int store, int L, LevelDev* lv, const float* r0, const uint32_t* /*mask*/,
                   Scal* sc, int coarse_f64, int visits, void** zfinal, KernelTimer* timer) {
    if (store == STORE_BF16) {
        if (coarse_f64) vcycle_t<__hip_bfloat16, __hip_bfloat16, double>(s, L, lv, r0, sc, zfinal, true, visits, timer);
        else vcycle_t<__hip_bfloat16, __hip_bfloat16, float>(s, L, lv, r0, sc, zfinal, true, visits, timer);
    } else {
        if (coarse_f64) vcycle_t<float, float, double>(s, L, lv, r0, sc, zfinal, false, visits, timer);
        else vcycle_t<float, float, float>(s, L, lv, r0, sc, zfinal, false, visits, timer);
    }
}
