// Conjugate-gradient kernels of the projection: the cached 7-point operator with fused
// dot products, the vector updates with fused norms, the mean-free residual projection and
// the FP64 refinement residual.  Each reduction kernel ends with a deterministic
// last-block fold that performs the scalar step of the reference's single-workgroup
// reduction kernels (alpha_reduce, beta_reduce, ...), so no separate launch is needed.
#include "mhdp_kernels.h"
#include "launch.h"
#include "rows.h"
#include "mob.h"

#define CONVERGED(sc) ((sc)->conv > 0.5)

// pressure_refresh (solver.wgsl:185-197) and pressure_check (solver.wgsl:317-323):
// t = b - L x on every cell (solid cells give 0, a pinned anchor its full physical row);
// r = t on active cells, 0 elsewhere; p = r/d.  Partials (r^2/d, t^2, b^2, r).
struct K_k_refresh {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const float* x;
        const float* b;
        const float* bx;
        const float* by;
        const float* bz;
        const float* dinv;
        float* r;
        float* p;
        double* part;
        Scal* sc;
        int meanfree;
        int check_only;
        int mode;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const float* __restrict__ x = A_.x;
        const float* __restrict__ b = A_.b;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ r = A_.r;
        float* __restrict__ p = A_.p;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int meanfree = A_.meanfree;
        int check_only = A_.check_only;
        int mode = A_.mode;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int cx, cy, cz;
            decode(g, i, cx, cy, cz);
            uint32_t nb[6];
            float w[6];
            stencil6(g, cx, cy, cz, i, bx, by, bz, nb, w);
            const float c = x[i];
            float a = 0.0f;
    #pragma unroll
            for (int k = 0; k < 6; k++) a += w[k] * (c - x[nb[k]]);
            const float bi = b[i];
            const float t = bi - a;
            const float di = dinv[i];
            const float ri = di != 0.0f ? t : 0.0f;
            if (!check_only) {
                r[i] = ri;
                p[i] = ri * di;
            }
            acc[0] += (double)ri * (double)ri * (double)di;
            acc[1] += (double)t * (double)t;
            acc[2] += (double)bi * (double)bi;
            acc[3] += meanfree ? (double)ri : 0.0;
        }
        reduce_finish<4>(vg, acc, part, sc, 0, mode);
    }
};

// project_pressure_residual (pressure_subspace.wgsl:8-24): r -= mean on active cells;
// initial form also restarts p = r/d.  Partials (r^2/d, r^2).
struct K_k_subspace {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float* r;
        float* p;
        const float* dinv;
        double* part;
        Scal* sc;
        int initial;
        int mode;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        float* __restrict__ r = A_.r;
        float* __restrict__ p = A_.p;
        const float* __restrict__ dinv = A_.dinv;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int initial = A_.initial;
        int mode = A_.mode;
        (void)vg;
        if (!initial && CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        const float mean = (float)sc->mean;
        double acc[2] = {0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const float di = dinv[i];
            if (di == 0.0f) continue;
            const float ri = r[i] - mean;
            r[i] = ri;
            if (initial) p[i] = ri * di;
            acc[0] += (double)ri * (double)ri * (double)di;
            acc[1] += (double)ri * (double)ri;
        }
        reduce_finish<2>(vg, acc, part, sc, 1, mode);
    }
};

// pressure_apply + alpha_reduce (solver.wgsl:204-215): q = A p, alpha = r.z / p.q.
struct K_k_apply {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const float* p;
        const float* bx;
        const float* by;
        const float* bz;
        float* q;
        double* part;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const float* __restrict__ p = A_.p;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        float* __restrict__ q = A_.q;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        double acc[1] = {0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int cx, cy, cz;
            decode(g, i, cx, cy, cz);
            uint32_t nb[6];
            float w[6];
            stencil6(g, cx, cy, cz, i, bx, by, bz, nb, w);
            const float c = p[i];
            float a = 0.0f;
    #pragma unroll
            for (int k = 0; k < 6; k++) a += w[k] * (c - p[nb[k]]);
            q[i] = a;
            acc[0] += (double)c * (double)a;
        }
        reduce_finish<1>(vg, acc, part, sc, 2, FIN_ALPHA);
    }
};

// update (solver.wgsl:216-230): x += alpha p (mean-free: alpha (p - p[anchor])), r -= alpha q.
// Partials (r^2/d, r^2, 0, r).
struct K_k_update {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float* x;
        float* r;
        const float* p;
        const float* q;
        const float* dinv;
        double* part;
        Scal* sc;
        int meanfree;
        int mode;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        float* __restrict__ x = A_.x;
        float* __restrict__ r = A_.r;
        const float* __restrict__ p = A_.p;
        const float* __restrict__ q = A_.q;
        const float* __restrict__ dinv = A_.dinv;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int meanfree = A_.meanfree;
        int mode = A_.mode;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        const float alpha = (float)sc->alpha;
        const float pa = meanfree ? p[sc->anchor] : 0.0f;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const float di = dinv[i];
            if (di == 0.0f) continue;
            const float ri = r[i] - alpha * q[i];
            x[i] = x[i] + alpha * (p[i] - pa);
            r[i] = ri;
            acc[0] += (double)ri * (double)ri * (double)di;
            acc[1] += (double)ri * (double)ri;
            acc[3] += (double)ri;
        }
        reduce_finish<4>(vg, acc, part, sc, 3, mode);
    }
};

// search (solver.wgsl:240-244): p = r/d + beta p.
struct K_k_search {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const float* r;
        const float* dinv;
        float* p;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const float* __restrict__ r = A_.r;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ p = A_.p;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        const float beta = (float)sc->beta;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const float di = dinv[i];
            if (di == 0.0f) continue;
            p[i] = r[i] * di + beta * p[i];
        }
    }
};

// dot_partial (multigrid.wgsl:90-100) + the following initial_reduce / beta_reduce.
template <class VT>
struct K_k_pdot {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const float* r;
        const VT* z;
        const float* dinv;
        float* p;
        double* part;
        Scal* sc;
        int initial;
        int mode;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const float* __restrict__ r = A_.r;
        const VT* __restrict__ z = A_.z;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ p = A_.p;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int initial = A_.initial;
        int mode = A_.mode;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        double acc[1] = {0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            if (dinv[i] == 0.0f) continue;
            const float zi = ldf(z[i]);
            acc[0] += (double)r[i] * (double)zi;
            if (initial) p[i] = zi;
        }
        reduce_finish<1>(vg, acc, part, sc, 4, mode);
    }
};

// preconditioned_search (multigrid.wgsl:102): p = z + beta p.
template <class VT>
struct K_k_psearch {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const VT* z;
        const float* dinv;
        float* p;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const VT* __restrict__ z = A_.z;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ p = A_.p;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        const float beta = (float)sc->beta;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            if (dinv[i] == 0.0f) continue;
            p[i] = ldf(z[i]) + beta * p[i];
        }
    }
};

// FP64 refinement residual: t = b - L x in FP64 (FP32 mobilities promoted) on every cell;
// the next inner FP32 right-hand side is r32 = t on active cells.  Partials (0, t^2, b^2, r).
struct K_k_residual64 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const double* x;
        const double* b;
        const float* bx;
        const float* by;
        const float* bz;
        const float* dinv;
        float* r32;
        double* part;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const double* __restrict__ x = A_.x;
        const double* __restrict__ b = A_.b;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ r32 = A_.r32;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int cx, cy, cz;
            decode(g, i, cx, cy, cz);
            uint32_t nb[6];
            float w[6];
            stencil6(g, cx, cy, cz, i, bx, by, bz, nb, w);
            const double c = x[i];
            double a = 0.0;
    #pragma unroll
            for (int k = 0; k < 6; k++) a += (double)w[k] * (c - x[nb[k]]);
            const double bi = b[i];
            const double t = bi - a;
            const bool active = dinv[i] != 0.0f;
            r32[i] = active ? (float)t : 0.0f;
            acc[1] += t * t;
            acc[2] += bi * bi;
            acc[3] += active ? t : 0.0;
        }
        reduce_finish<4>(vg, acc, part, sc, 5, FIN_NORM64);
    }
};

struct K_k_xupdate64 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        double* x;
        const float* e;
        const float* dinv;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        double* __restrict__ x = A_.x;
        const float* __restrict__ e = A_.e;
        const float* __restrict__ dinv = A_.dinv;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK)
            if (dinv[i] != 0.0f) x[i] += (double)e[i];
    }
};

// One FP64 refinement step's start (fused multigrid path): t = b - L x in FP64 (x = 0 on
// the first step: t = b), r32 = t on active cells, and the inner solve's x, p, p' zeroed.
// Partials (sum r32^2, sum t^2, sum b^2, sum r32) -> FIN_OUTER.
struct K_k_outer_start {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const double* x;
        const double* b;
        const float* bx;
        const float* by;
        const float* bz;
        const float* dinv;
        float* r32;
        float* x32;
        float* p0;
        float* p1;
        double* part;
        Scal* sc;
        int zero_x;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const double* __restrict__ x = A_.x;
        const double* __restrict__ b = A_.b;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ r32 = A_.r32;
        float* __restrict__ x32 = A_.x32;
        float* __restrict__ p0 = A_.p0;
        float* __restrict__ p1 = A_.p1;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int zero_x = A_.zero_x;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const double bi = b[i];
            double t = bi;
            if (!zero_x) {
                int cx, cy, cz;
                decode(g, i, cx, cy, cz);
                uint32_t nb[6];
                float w[6];
                stencil6(g, cx, cy, cz, i, bx, by, bz, nb, w);
                const double c = x[i];
                double a = 0.0;
    #pragma unroll
                for (int k = 0; k < 6; k++) a += (double)w[k] * (c - x[nb[k]]);
                t = bi - a;
            }
            const float rv = dinv[i] != 0.0f ? (float)t : 0.0f;
            r32[i] = rv;
            x32[i] = 0.0f;
            if (p0) {  // null: the brick apply does not read the old search direction at a start
                p0[i] = 0.0f;
                p1[i] = 0.0f;
            }
            acc[0] += (double)rv * (double)rv;
            acc[1] += t * t;
            acc[2] += bi * bi;
            acc[3] += (double)rv;
        }
        reduce_finish<4>(vg, acc, part, sc, 5, FIN_OUTER);
    }
};

// x64 += e on active cells, e = the inner FP32 solution with the fused path's owed update
// x += alpha (p_k - p_k[anchor]) applied first (same FP32 arithmetic as the flush);
// `assign`: x64 = e (x64 was zero), inactive cells 0.
struct K_k_outer_update {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        double* x64;
        const float* x;
        const float* p0;
        const float* p1;
        const float* dinv;
        const Scal* sc;
        int meanfree;
        int assign;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        double* __restrict__ x64 = A_.x64;
        const float* __restrict__ x = A_.x;
        const float* p0 = A_.p0;
        const float* p1 = A_.p1;
        const float* __restrict__ dinv = A_.dinv;
        const Scal* sc = A_.sc;
        int meanfree = A_.meanfree;
        int assign = A_.assign;
        (void)vg;
        const bool pending = sc->pending != 0u;
        const uint32_t it = (uint32_t)sc->iters;
        const float* __restrict__ pk = ((it - 1u) & 1u) ? p1 : p0;
        const float alpha = (float)sc->alpha;
        const float pa = (pending && meanfree) ? pk[sc->anchor] : 0.0f;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            if (dinv[i] != 0.0f) {
                float e = x[i];
                if (pending) e = e + alpha * (pk[i] - pa);
                x64[i] = assign ? (double)e : x64[i] + (double)e;
            } else if (assign) {
                x64[i] = 0.0;
            }
        }
    }
};

// Row-vector forms of the two refinement passes (nx % 4 == 0): four cells of a row per thread,
// the same FP64 arithmetic per cell (stencil order x-, x+, y-, y+, z-, z+).
struct K_k_outer_start4 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const double* x;
        const double* b;
        const float* bx;
        const float* by;
        const float* bz;
        const float* dinv;
        float* r32;
        float* x32;
        float* p0;
        float* p1;
        double* part;
        Scal* sc;
        int zero_x;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const double* __restrict__ x = A_.x;
        const double* __restrict__ b = A_.b;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ r32 = A_.r32;
        float* __restrict__ x32 = A_.x32;
        float* __restrict__ p0 = A_.p0;
        float* __restrict__ p1 = A_.p1;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int zero_x = A_.zero_x;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n >> 2, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const float zero4[4] = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t t = lo + threadIdx.x; t < hi; t += MHDP_BLOCK) {
            const Row4 r = row4(g, t);
            double bv[4], tv[4];
            ld4d(b, r.i0, bv);
            float di[4];
            ld4f(dinv, r.i0, di);
    #pragma unroll
            for (int k = 0; k < 4; k++) tv[k] = bv[k];
            if (!zero_x) {
                double xc[4], xym[4], xyp[4], xzm[4], xzp[4];
                ld4d(x, r.i0, xc);
                ld4d(x, r.ym, xym);
                ld4d(x, r.yp, xyp);
                ld4d(x, r.zm, xzm);
                ld4d(x, r.zp, xzp);
                const double xxm = x[r.xm], xxp = x[r.xp];
                float wx[4], wy[4], wz[4], wyu[4], wzu[4];
                ld4f(bx, r.i0, wx);
                ld4f(by, r.i0, wy);
                ld4f(bz, r.i0, wz);
                ld4f(by, r.yp, wyu);
                ld4f(bz, r.zp, wzu);
                const float wxe = bx[r.xp];
    #pragma unroll
                for (int k = 0; k < 4; k++) {
                    const double c = xc[k];
                    const double w1 = k < 3 ? (double)wx[k + 1] : (r.xp_ok ? (double)wxe : 0.0);
                    double a = 0.0;
                    a += (double)wx[k] * (c - (k > 0 ? xc[k - 1] : xxm));
                    a += w1 * (c - (k < 3 ? xc[k + 1] : xxp));
                    a += (double)wy[k] * (c - xym[k]);
                    a += (r.yp_ok ? (double)wyu[k] : 0.0) * (c - xyp[k]);
                    a += (double)wz[k] * (c - xzm[k]);
                    a += (r.zp_ok ? (double)wzu[k] : 0.0) * (c - xzp[k]);
                    tv[k] = bv[k] - a;
                }
            }
            float rv[4];
    #pragma unroll
            for (int k = 0; k < 4; k++) {
                rv[k] = di[k] != 0.0f ? (float)tv[k] : 0.0f;
                acc[0] += (double)rv[k] * (double)rv[k];
                acc[1] += tv[k] * tv[k];
                acc[2] += bv[k] * bv[k];
                acc[3] += (double)rv[k];
            }
            st4f(r32, r.i0, rv);
            st4f(x32, r.i0, zero4);
            if (p0) {  // null: the brick apply does not read the old search direction at a start
                st4f(p0, r.i0, zero4);
                st4f(p1, r.i0, zero4);
            }
        }
        reduce_finish<4>(vg, acc, part, sc, 5, FIN_OUTER);
    }
};

// K_k_outer_start4 with the mobilities formed from the liquid fraction and the open bytes (mob.h)
// instead of loaded, and the active cells from the open bytes: the same residual bit for bit.
template <bool TWO>
struct K_k_outer_start4m {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const double* x;
        const double* b;
        const float* alpha;  // TWO
        const uint8_t* open8;
        MobP mp;
        int meanfree;
        float* r32;
        float* x32;
        float* p0;
        float* p1;
        double* part;
        Scal* sc;
        int zero_x;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        const Grid3 g = A_.g;
        const double* __restrict__ x = A_.x;
        const uint8_t* __restrict__ open8 = A_.open8;
        const float* __restrict__ al = A_.alpha;
        const MobP mp = A_.mp;
        uint32_t lo, hi;
        block_range(vg, g.n >> 2, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const float zero4[4] = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t t = lo + threadIdx.x; t < hi; t += MHDP_BLOCK) {
            const Row4 r = row4(g, t);
            double bv[4], tv[4];
            ld4d(A_.b, r.i0, bv);
            const uint32_t ow = *reinterpret_cast<const uint32_t*>(open8 + r.i0);
#pragma unroll
            for (int k = 0; k < 4; k++) tv[k] = bv[k];
            if (!A_.zero_x) {
                double xc[4], xym[4], xyp[4], xzm[4], xzp[4];
                ld4d(x, r.i0, xc);
                ld4d(x, r.ym, xym);
                ld4d(x, r.yp, xyp);
                ld4d(x, r.zm, xzm);
                ld4d(x, r.zp, xzp);
                const double xxm = x[r.xm], xxp = x[r.xp];
                const uint32_t oyu = *reinterpret_cast<const uint32_t*>(open8 + r.yp);
                const uint32_t ozu = *reinterpret_cast<const uint32_t*>(open8 + r.zp);
                const uint32_t oxu = open8[r.xp];
                float dc[4] = {0.f, 0.f, 0.f, 0.f}, dym[4] = {0.f, 0.f, 0.f, 0.f}, dyp[4] = {0.f, 0.f, 0.f, 0.f};
                float dzm[4] = {0.f, 0.f, 0.f, 0.f}, dzp[4] = {0.f, 0.f, 0.f, 0.f}, dxm = 0.f, dxp = 0.f;
                if (TWO) {
                    float a[4];
                    ld4f(al, r.i0, a);
#pragma unroll
                    for (int k = 0; k < 4; k++) dc[k] = mob_dens(mp, a[k]);
                    ld4f(al, r.ym, a);
#pragma unroll
                    for (int k = 0; k < 4; k++) dym[k] = mob_dens(mp, a[k]);
                    ld4f(al, r.yp, a);
#pragma unroll
                    for (int k = 0; k < 4; k++) dyp[k] = mob_dens(mp, a[k]);
                    ld4f(al, r.zm, a);
#pragma unroll
                    for (int k = 0; k < 4; k++) dzm[k] = mob_dens(mp, a[k]);
                    ld4f(al, r.zp, a);
#pragma unroll
                    for (int k = 0; k < 4; k++) dzp[k] = mob_dens(mp, a[k]);
                    dxm = mob_dens(mp, al[r.xm]);
                    dxp = mob_dens(mp, al[r.xp]);
                }
                float wx[5];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    wx[k] = mob_face<TWO>(mp, obyte(ow, k) & 1u, k > 0 ? dc[k - 1] : dxm, dc[k]);
                wx[4] = mob_face<TWO>(mp, r.xp_ok && (oxu & 1u), dc[3], dxp);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t o = obyte(ow, k);
                    const float wy = mob_face<TWO>(mp, o & 2u, dym[k], dc[k]);
                    const float wyu = mob_face<TWO>(mp, r.yp_ok && (obyte(oyu, k) & 2u), dc[k], dyp[k]);
                    const float wz = mob_face<TWO>(mp, o & 4u, dzm[k], dc[k]);
                    const float wzu = mob_face<TWO>(mp, r.zp_ok && (obyte(ozu, k) & 4u), dc[k], dzp[k]);
                    const double c = xc[k];
                    double a = 0.0;
                    a += (double)wx[k] * (c - (k > 0 ? xc[k - 1] : xxm));
                    a += (double)wx[k + 1] * (c - (k < 3 ? xc[k + 1] : xxp));
                    a += (double)wy * (c - xym[k]);
                    a += (double)wyu * (c - xyp[k]);
                    a += (double)wz * (c - xzm[k]);
                    a += (double)wzu * (c - xzp[k]);
                    tv[k] = bv[k] - a;
                }
            }
            float rv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                rv[k] = open_active(obyte(ow, k), A_.meanfree) ? (float)tv[k] : 0.0f;
                acc[0] += (double)rv[k] * (double)rv[k];
                acc[1] += tv[k] * tv[k];
                acc[2] += bv[k] * bv[k];
                acc[3] += (double)rv[k];
            }
            st4f(A_.r32, r.i0, rv);
            st4f(A_.x32, r.i0, zero4);
            if (A_.p0) {  // null: the brick apply does not read the old search direction at a start
                st4f(A_.p0, r.i0, zero4);
                st4f(A_.p1, r.i0, zero4);
            }
        }
        reduce_finish<4>(vg, acc, A_.part, A_.sc, 5, FIN_OUTER);
    }
};

struct K_k_outer_update4 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        double* x64;
        const float* x;
        const float* p0;
        const float* p1;
        const float* dinv;
        const Scal* sc;
        int meanfree;
        int assign;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        double* __restrict__ x64 = A_.x64;
        const float* __restrict__ x = A_.x;
        const float* p0 = A_.p0;
        const float* p1 = A_.p1;
        const float* __restrict__ dinv = A_.dinv;
        const Scal* sc = A_.sc;
        int meanfree = A_.meanfree;
        int assign = A_.assign;
        (void)vg;
        const bool pending = sc->pending != 0u;
        const uint32_t it = (uint32_t)sc->iters;
        const float* __restrict__ pk = ((it - 1u) & 1u) ? p1 : p0;
        const float alpha = (float)sc->alpha;
        const float pa = (pending && meanfree) ? pk[sc->anchor] : 0.0f;
        uint32_t lo, hi;
        block_range(vg, g.n >> 2, lo, hi);
        for (uint32_t t = lo + threadIdx.x; t < hi; t += MHDP_BLOCK) {
            const uint32_t i0 = t << 2;
            float xv[4], pv[4], di[4];
            ld4f(x, i0, xv);
            ld4f(pk, i0, pv);
            ld4f(dinv, i0, di);
            double xo[4] = {0.0, 0.0, 0.0, 0.0};
            if (!assign) ld4d(x64, i0, xo);
    #pragma unroll
            for (int k = 0; k < 4; k++) {
                float e = xv[k];
                if (pending) e = e + alpha * (pv[k] - pa);
                if (di[k] != 0.0f) xo[k] = assign ? (double)e : xo[k] + (double)e;
            }
            st4d(x64, i0, xo);
        }
    }
};

template <class T>
struct K_k_fill {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        T* dst;
        T v;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        uint32_t n = A_.n;
        T* __restrict__ dst = A_.dst;
        T v = A_.v;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) dst[i] = v;
    }
};

struct K_k_cast_f32_f64 {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const float* s;
        double* d;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        uint32_t n = A_.n;
        const float* __restrict__ s = A_.s;
        double* __restrict__ d = A_.d;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) d[i] = (double)s[i];
    }
};

// Plain operator application (test hook): q = sum beta (p_c - p_n) on every cell; with
// identity_inactive, inactive cells return p (the pinned-anchor identity row, solver.wgsl:79).
struct K_k_apply_plain {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const float* p;
        const float* bx;
        const float* by;
        const float* bz;
        const float* dinv;
        int identity_inactive;
        float* q;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const float* __restrict__ p = A_.p;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const float* __restrict__ dinv = A_.dinv;
        int identity_inactive = A_.identity_inactive;
        float* __restrict__ q = A_.q;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int cx, cy, cz;
            decode(g, i, cx, cy, cz);
            uint32_t nb[6];
            float w[6];
            stencil6(g, cx, cy, cz, i, bx, by, bz, nb, w);
            const float c = p[i];
            float a = 0.0f;
    #pragma unroll
            for (int k = 0; k < 6; k++) a += w[k] * (c - p[nb[k]]);
            q[i] = (identity_inactive && dinv[i] == 0.0f) ? c : a;
        }
    }
};

// ---------------------------------------------------------------- launchers

void launch_refresh(hipStream_t s, const Grid3& g, const float* x, const float* b, const float* bx, const float* by,
                    const float* bz, const float* dinv, float* r, float* p, double* part, Scal* sc, int meanfree,
                    int check_only, int mode) {
    go<K_k_refresh>(s, grid_blocks(g.n), {g, x, b, bx, by, bz, dinv, r, p, part, sc, meanfree, check_only, mode});
}
void launch_subspace(hipStream_t s, const Grid3& g, float* r, float* p, const float* dinv, double* part, Scal* sc,
                     int initial, int mode) {
    go<K_k_subspace>(s, grid_blocks(g.n), {g, r, p, dinv, part, sc, initial, mode});
}
void launch_apply(hipStream_t s, const Grid3& g, const float* p, const float* bx, const float* by, const float* bz,
                  float* q, double* part, Scal* sc) {
    go<K_k_apply>(s, grid_blocks(g.n), {g, p, bx, by, bz, q, part, sc});
}
void launch_update(hipStream_t s, const Grid3& g, float* x, float* r, const float* p, const float* q,
                   const float* dinv, double* part, Scal* sc, int meanfree, int mode) {
    go<K_k_update>(s, grid_blocks(g.n), {g, x, r, p, q, dinv, part, sc, meanfree, mode});
}
void launch_search(hipStream_t s, const Grid3& g, const float* r, const float* dinv, float* p, Scal* sc) {
    go<K_k_search>(s, grid_blocks(g.n), {g, r, dinv, p, sc});
}
void launch_pdot(hipStream_t s, const Grid3& g, int store, const float* r, const void* z, const float* dinv, float* p,
                 double* part, Scal* sc, int initial, int mode) {
    if (store == STORE_BF16)
        go<K_k_pdot<__hip_bfloat16>>(s, grid_blocks(g.n), {g, r, (const __hip_bfloat16*)z, dinv, p, part, sc, initial, mode});
    else
        go<K_k_pdot<float>>(s, grid_blocks(g.n), {g, r, (const float*)z, dinv, p, part, sc, initial, mode});
}
void launch_psearch(hipStream_t s, const Grid3& g, int store, const void* z, const float* dinv, float* p, Scal* sc) {
    if (store == STORE_BF16)
        go<K_k_psearch<__hip_bfloat16>>(s, grid_blocks(g.n), {g, (const __hip_bfloat16*)z, dinv, p, sc});
    else
        go<K_k_psearch<float>>(s, grid_blocks(g.n), {g, (const float*)z, dinv, p, sc});
}
void launch_residual64(hipStream_t s, const Grid3& g, const double* x64, const double* b64, const float* bx,
                       const float* by, const float* bz, const float* dinv, float* r32, double* part, Scal* sc) {
    go<K_k_residual64>(s, grid_blocks(g.n), {g, x64, b64, bx, by, bz, dinv, r32, part, sc});
}
void launch_outer_start(hipStream_t s, const Grid3& g, const double* x64, const double* b64, const float* bx,
                        const float* by, const float* bz, const float* dinv, float* r32, float* x32, float* p0,
                        float* p1, double* part, Scal* sc, bool zero_x) {
    if ((g.nx & 3) == 0)
        go<K_k_outer_start4>(s, grid_blocks(g.n >> 2), {g, x64, b64, bx, by, bz, dinv, r32, x32, p0, p1, part, sc, zero_x ? 1 : 0});
    else
        go<K_k_outer_start>(s, grid_blocks(g.n), {g, x64, b64, bx, by, bz, dinv, r32, x32, p0, p1, part, sc, zero_x ? 1 : 0});
}
bool launch_outer_start_m(hipStream_t s, const Grid3& g, const double* x64, const double* b64, const float* alpha,
                          const uint8_t* open8, float rho_l, float rho_air, bool meanfree, float* r32, float* x32,
                          float* p0, float* p1, double* part, Scal* sc, bool zero_x) {
    if ((g.nx & 3) != 0 || !open8) return false;
    const MobP mp{rho_l, rho_air};
    if (alpha)
        go<K_k_outer_start4m<true>>(s, grid_blocks(g.n >> 2), {g, x64, b64, alpha, open8, mp, meanfree ? 1 : 0, r32, x32,
                                                               p0, p1, part, sc, zero_x ? 1 : 0});
    else
        go<K_k_outer_start4m<false>>(s, grid_blocks(g.n >> 2), {g, x64, b64, alpha, open8, mp, meanfree ? 1 : 0, r32, x32,
                                                                p0, p1, part, sc, zero_x ? 1 : 0});
    return true;
}
void launch_outer_update(hipStream_t s, const Grid3& g, double* x64, const float* x, const float* p0, const float* p1,
                         const float* dinv, const Scal* sc, bool meanfree, bool assign) {
    if ((g.nx & 3) == 0)
        go<K_k_outer_update4>(s, grid_blocks(g.n >> 2), {g, x64, x, p0, p1, dinv, sc, meanfree ? 1 : 0, assign ? 1 : 0});
    else
        go<K_k_outer_update>(s, grid_blocks(g.n), {g, x64, x, p0, p1, dinv, sc, meanfree ? 1 : 0, assign ? 1 : 0});
}
void launch_xupdate64(hipStream_t s, const Grid3& g, double* x64, const float* e, const float* dinv) {
    go<K_k_xupdate64>(s, grid_blocks(g.n), {g, x64, e, dinv});
}
void launch_fill_f32(hipStream_t s, uint32_t n, float* dst, float v) {
    go<K_k_fill<float>>(s, grid_blocks(n), {n, dst, v});
}
void launch_fill_f64(hipStream_t s, uint32_t n, double* dst, double v) {
    go<K_k_fill<double>>(s, grid_blocks(n), {n, dst, v});
}
void launch_apply_plain(hipStream_t s, const Grid3& g, const float* p, const float* bx, const float* by,
                        const float* bz, const float* dinv, int identity_inactive, float* q) {
    go<K_k_apply_plain>(s, grid_blocks(g.n), {g, p, bx, by, bz, dinv, identity_inactive, q});
}
void launch_cast_f32_f64(hipStream_t s, uint32_t n, const float* src, double* dst) {
    go<K_k_cast_f32_f64>(s, grid_blocks(n), {n, src, dst});
}
