// Fused multigrid-preconditioned CG iteration for the fine grid (the hot path).
//
// The reference runs each MG-PCG iteration as ~5L+4 dispatches (pressure_apply,
// alpha_reduce, update, residual mean/subspace, the V-cycle stages, preconditioned_dot,
// beta_reduce, preconditioned_search; multigrid.rs:302-326).  Here the fine-grid work of
// one iteration is four streaming passes over HBM:
//   1. k_fused_apply   p_k = z_k + beta p_{k-1} (ping-pong buffers), the deferred update
//                      x += alpha_{k-1} p_{k-1}, q = A p_k and p.q -> alpha
//   2. k_fused_update  r = (r - m) - alpha q with sum r, sum r^2 -> mean m and |r - m|^2
//   3. k_fused_restrict  pre-smooth z1 = omega (r - m)/d, residual r - A z1 and the sum
//                      over each 2x2x2 aggregate, in one pass (z1 is never stored)
//   4. k_fused_smooth  prolongation z2 = z1 + z_coarse, post-smooth
//                      z = z2 + omega (r - m - A z2)/d and (r - m).z -> beta
// The mean-free residual projection (pressure_subspace.wgsl) is deferred: passes 3, 4 and
// the next pass 2 read r - m.  The algebra is the reference's; only the order of the
// x update and the place of the mean subtraction change.
//
// Vector path: each thread owns V = 4 consecutive x cells (nx % 4 == 0), so FP32 arrays
// move as 16-byte and BF16 arrays as 8-byte loads; V = 1 serves any other grid.
#include "mhdp_kernels.h"
#include "launch.h"

#define CONVERGED(sc) ((sc)->conv > 0.5)

namespace {

template <int V>
struct VF {
    float v[V];
};

__device__ __forceinline__ float bf_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf_hi(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

template <int V>
__device__ __forceinline__ VF<V> ldv(const float* __restrict__ p, uint32_t i) {
    VF<V> o;
    if constexpr (V == 4) {
        const float4 a = *reinterpret_cast<const float4*>(p + i);
        o.v[0] = a.x; o.v[1] = a.y; o.v[2] = a.z; o.v[3] = a.w;
    } else {
#pragma unroll
        for (int k = 0; k < V; k++) o.v[k] = p[i + k];
    }
    return o;
}
template <int V>
__device__ __forceinline__ VF<V> ldv(const __hip_bfloat16* __restrict__ p, uint32_t i) {
    VF<V> o;
    if constexpr (V == 4) {
        const uint2 a = *reinterpret_cast<const uint2*>(p + i);
        o.v[0] = bf_lo(a.x); o.v[1] = bf_hi(a.x); o.v[2] = bf_lo(a.y); o.v[3] = bf_hi(a.y);
    } else {
#pragma unroll
        for (int k = 0; k < V; k++) o.v[k] = __bfloat162float(p[i + k]);
    }
    return o;
}
template <int V>
__device__ __forceinline__ void stv(float* __restrict__ p, uint32_t i, const VF<V>& a) {
    if constexpr (V == 4) {
        *reinterpret_cast<float4*>(p + i) = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]);
    } else {
#pragma unroll
        for (int k = 0; k < V; k++) p[i + k] = a.v[k];
    }
}
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    const __hip_bfloat16 x = __float2bfloat16(a), y = __float2bfloat16(b);
    return (uint32_t)__bfloat16_as_ushort(x) | ((uint32_t)__bfloat16_as_ushort(y) << 16);
}
template <int V>
__device__ __forceinline__ void stv(__hip_bfloat16* __restrict__ p, uint32_t i, const VF<V>& a) {
    if constexpr (V == 4) {
        *reinterpret_cast<uint2*>(p + i) = make_uint2(pack_bf16(a.v[0], a.v[1]), pack_bf16(a.v[2], a.v[3]));
    } else {
#pragma unroll
        for (int k = 0; k < V; k++) p[i + k] = __float2bfloat16(a.v[k]);
    }
}

// Neighbour geometry of a group of V cells starting at (x0, y, z), flat index i0.
struct Nb {
    uint32_t xm, xp;            // scalar neighbours left of the first / right of the last cell
    uint32_t ym, yp, zm, zp;    // group starts of the y/z neighbour rows
    int cxm, cxp, cym, cyp, czm, czp;  // their coordinates along the moved axis
    bool xp_open_ok, yp_ok, zp_ok;     // upper-face weights exist (else 0)
};
// Indices always stay in range: a closed boundary gets the cell itself (weight 0).
template <int V>
__device__ __forceinline__ Nb neighbours(const Grid3& g, int x0, int y, int z, uint32_t i0) {
    Nb n;
    const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
    if (x0 > 0) { n.xm = i0 - 1; n.cxm = x0 - 1; }
    else if (g.per & 1u) { n.xm = i0 + (uint32_t)(g.nx - 1); n.cxm = g.nx - 1; }
    else { n.xm = i0; n.cxm = x0; }
    const int xl = x0 + V - 1;
    if (xl + 1 < g.nx) { n.xp = i0 + V; n.cxp = xl + 1; n.xp_open_ok = true; }
    else if (g.per & 1u) { n.xp = i0 + V - (uint32_t)g.nx; n.cxp = 0; n.xp_open_ok = true; }
    else { n.xp = i0 + V - 1; n.cxp = xl; n.xp_open_ok = false; }
    if (y > 0) { n.ym = i0 - sy; n.cym = y - 1; }
    else if (g.per & 2u) { n.ym = i0 + (uint32_t)(g.ny - 1) * sy; n.cym = g.ny - 1; }
    else { n.ym = i0; n.cym = y; }
    if (y + 1 < g.ny) { n.yp = i0 + sy; n.cyp = y + 1; n.yp_ok = true; }
    else if (g.per & 2u) { n.yp = i0 - (uint32_t)(g.ny - 1) * sy; n.cyp = 0; n.yp_ok = true; }
    else { n.yp = i0; n.cyp = y; n.yp_ok = false; }
    if (z > 0) { n.zm = i0 - sz; n.czm = z - 1; }
    else if (g.per & 4u) { n.zm = i0 + (uint32_t)(g.nz - 1) * sz; n.czm = g.nz - 1; }
    else { n.zm = i0; n.czm = z; }
    if (z + 1 < g.nz) { n.zp = i0 + sz; n.czp = z + 1; n.zp_ok = true; }
    else if (g.per & 4u) { n.zp = i0 - (uint32_t)(g.nz - 1) * sz; n.czp = 0; n.zp_ok = true; }
    else { n.zp = i0; n.czp = z; n.zp_ok = false; }
    return n;
}

// Face weights of the group in the reference order x-, x+, y-, y+, z-, z+.
template <int V, class CT>
struct W6 {
    float w[6][V];
};
template <int V, class CT>
__device__ __forceinline__ W6<V, CT> weights(const Nb& n, uint32_t i0, const CT* __restrict__ bx,
                                             const CT* __restrict__ by, const CT* __restrict__ bz) {
    W6<V, CT> W;
    const VF<V> lx = ldv<V>(bx, i0);
    const float lastx = n.xp_open_ok ? ldf(bx[n.xp]) : 0.0f;
    const VF<V> ly = ldv<V>(by, i0), lz = ldv<V>(bz, i0);
    VF<V> uy, uz;
    if (n.yp_ok) uy = ldv<V>(by, n.yp);
    else
#pragma unroll
        for (int k = 0; k < V; k++) uy.v[k] = 0.0f;
    if (n.zp_ok) uz = ldv<V>(bz, n.zp);
    else
#pragma unroll
        for (int k = 0; k < V; k++) uz.v[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < V; k++) {
        W.w[0][k] = lx.v[k];
        W.w[1][k] = k + 1 < V ? lx.v[k + 1] : lastx;
        W.w[2][k] = ly.v[k];
        W.w[3][k] = uy.v[k];
        W.w[4][k] = lz.v[k];
        W.w[5][k] = uz.v[k];
    }
    return W;
}

// Values of a field at the group and its six neighbour positions.
template <int V>
struct Star {
    float c[V], xm[V], xp[V], ym[V], yp[V], zm[V], zp[V];
};

template <int V, class F>
__device__ __forceinline__ Star<V> gather(const Nb& n, uint32_t i0, F&& field_v, float xm_scalar, float xp_scalar) {
    Star<V> s;
    const VF<V> c = field_v(i0), ym = field_v(n.ym), yp = field_v(n.yp), zm = field_v(n.zm), zp = field_v(n.zp);
#pragma unroll
    for (int k = 0; k < V; k++) {
        s.c[k] = c.v[k];
        s.ym[k] = ym.v[k];
        s.yp[k] = yp.v[k];
        s.zm[k] = zm.v[k];
        s.zp[k] = zp.v[k];
        s.xm[k] = k > 0 ? c.v[k - 1] : xm_scalar;
        s.xp[k] = k + 1 < V ? c.v[k + 1] : xp_scalar;
    }
    return s;
}

template <int V, class CT>
__device__ __forceinline__ float star_action(const W6<V, CT>& W, const Star<V>& s, int k) {
    float a = 0.0f;
    a += W.w[0][k] * (s.c[k] - s.xm[k]);
    a += W.w[1][k] * (s.c[k] - s.xp[k]);
    a += W.w[2][k] * (s.c[k] - s.ym[k]);
    a += W.w[3][k] * (s.c[k] - s.yp[k]);
    a += W.w[4][k] * (s.c[k] - s.zm[k]);
    a += W.w[5][k] * (s.c[k] - s.zp[k]);
    return a;
}

}  // namespace

// ---------------------------------------------------------------- pass 1
// p_k = z_k + beta p_{k-1}; x += alpha_{k-1} (p_{k-1} - p_{k-1}[anchor]) if pending;
// q = A p_k (FP32 operator); partial p.q -> FIN_ALPHA.  p buffers alternate with the
// iteration count held on the device (iters even: write p0).
template <int V, class VT>
struct K_k_fused_apply {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const VT* z;
        float* p0;
        float* p1;
        float* x;
        const float* bx;
        const float* by;
        const float* bz;
        const float* dinv;
        float* q;
        double* part;
        Scal* sc;
        int meanfree;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const VT* __restrict__ z = A_.z;
        float* p0 = A_.p0;
        float* p1 = A_.p1;
        float* __restrict__ x = A_.x;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const float* __restrict__ dinv = A_.dinv;
        float* __restrict__ q = A_.q;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int meanfree = A_.meanfree;
        (void)vg;
        if (CONVERGED(sc)) return;
        const uint32_t it = (uint32_t)sc->iters;
        const float* __restrict__ pold = (it & 1u) ? p0 : p1;
        float* __restrict__ pnew = (it & 1u) ? p1 : p0;
        const float beta = (float)sc->beta;
        const bool pending = sc->pending != 0;
        const float alpha = (float)sc->alpha;
        const float pa = (pending && meanfree) ? pold[sc->anchor] : 0.0f;
        uint32_t lo, hi;
        block_range(vg, g.n / V, lo, hi);
        double acc[1] = {0.0};
        for (uint32_t gi = lo + threadIdx.x; gi < hi; gi += MHDP_BLOCK) {
            const uint32_t i0 = gi * V;
            int x0, y, zz;
            decode(g, i0, x0, y, zz);
            const Nb n = neighbours<V>(g, x0, y, zz, i0);
            const W6<V, float> W = weights<V, float>(n, i0, bx, by, bz);
            auto pv = [&](uint32_t j) {
                const VF<V> a = ldv<V>(z, j), b = ldv<V>(pold, j);
                VF<V> o;
    #pragma unroll
                for (int k = 0; k < V; k++) o.v[k] = a.v[k] + beta * b.v[k];
                return o;
            };
            const float pxm = ldf(z[n.xm]) + beta * pold[n.xm];
            const float pxp = ldf(z[n.xp]) + beta * pold[n.xp];
            const Star<V> s = gather<V>(n, i0, pv, pxm, pxp);
            const VF<V> di = ldv<V>(dinv, i0);
            VF<V> qv, pn;
    #pragma unroll
            for (int k = 0; k < V; k++) {
                const bool act = di.v[k] != 0.0f;
                pn.v[k] = act ? s.c[k] : 0.0f;
                qv.v[k] = star_action<V, float>(W, s, k);
                acc[0] += (double)pn.v[k] * (double)qv.v[k];
            }
            stv<V>(pnew, i0, pn);
            stv<V>(q, i0, qv);
            if (pending) {
                const VF<V> po = ldv<V>(pold, i0);
                VF<V> xv = ldv<V>(x, i0);
    #pragma unroll
                for (int k = 0; k < V; k++)
                    if (di.v[k] != 0.0f) xv.v[k] = xv.v[k] + alpha * (po.v[k] - pa);
                stv<V>(x, i0, xv);
            }
        }
        reduce_finish<1>(vg, acc, part, sc, 2, FIN_ALPHA);
    }
};

// Apply the x update still pending from the last iteration (before a residual check,
// a refresh or the end of the solve).
template <int V>
struct K_k_fused_xflush {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        const float* p0;
        const float* p1;
        float* x;
        const float* dinv;
        Scal* sc;
        int meanfree;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        const float* p0 = A_.p0;
        const float* p1 = A_.p1;
        float* __restrict__ x = A_.x;
        const float* __restrict__ dinv = A_.dinv;
        Scal* sc = A_.sc;
        int meanfree = A_.meanfree;
        (void)vg;
        if (sc->pending == 0) return;
        const uint32_t it = (uint32_t)sc->iters;
        // the last completed iteration k = iters-1 wrote p_k to p0 when k is even
        const float* __restrict__ pk = ((it - 1u) & 1u) ? p1 : p0;
        const float alpha = (float)sc->alpha;
        const float pa = meanfree ? pk[sc->anchor] : 0.0f;
        uint32_t lo, hi;
        block_range(vg, g.n / V, lo, hi);
        for (uint32_t gi = lo + threadIdx.x; gi < hi; gi += MHDP_BLOCK) {
            const uint32_t i0 = gi * V;
            const VF<V> po = ldv<V>(pk, i0), di = ldv<V>(dinv, i0);
            VF<V> xv = ldv<V>(x, i0);
    #pragma unroll
            for (int k = 0; k < V; k++)
                if (di.v[k] != 0.0f) xv.v[k] = xv.v[k] + alpha * (po.v[k] - pa);
            stv<V>(x, i0, xv);
        }
    }
};

// ---------------------------------------------------------------- pass 2
// r = (r - m) - alpha q on active cells; partials (0, sum r^2, 0, sum r) -> FIN_UPDATE_MG.
template <int V>
struct K_k_fused_update {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float* r;
        const float* q;
        const float* dinv;
        double* part;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        float* __restrict__ r = A_.r;
        const float* __restrict__ q = A_.q;
        const float* __restrict__ dinv = A_.dinv;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float alpha = (float)sc->alpha;
        const float m = (float)sc->mean;
        uint32_t lo, hi;
        block_range(vg, g.n / V, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t gi = lo + threadIdx.x; gi < hi; gi += MHDP_BLOCK) {
            const uint32_t i0 = gi * V;
            const VF<V> di = ldv<V>(dinv, i0), qv = ldv<V>(q, i0);
            VF<V> rv = ldv<V>(r, i0);
    #pragma unroll
            for (int k = 0; k < V; k++) {
                const bool act = di.v[k] != 0.0f;
                const float rn = act ? (rv.v[k] - m) - alpha * qv.v[k] : 0.0f;
                rv.v[k] = rn;
                acc[1] += (double)rn * (double)rn;
                acc[3] += (double)rn;
            }
            stv<V>(r, i0, rv);
        }
        reduce_finish<4>(vg, acc, part, sc, 3, FIN_UPDATE_MG);
    }
};

// ---------------------------------------------------------------- pass 3
// For each coarse cell: sum over its active children of (r - m) - A z1, z1 = omega (r-m)/d.
template <class CT, class VT>
struct K_k_fused_restrict {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 cg;
        Grid3 fg;
        const float* r;
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
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
        VT* __restrict__ rc = A_.rc;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
        const float m = (float)sc->mean;
        uint32_t lo, hi;
        block_range(vg, cg.n, lo, hi);
        const uint32_t sy = (uint32_t)fg.nx, sz = (uint32_t)fg.nx * (uint32_t)fg.ny;
        for (uint32_t ci = lo + threadIdx.x; ci < hi; ci += MHDP_BLOCK) {
            int qx, qy, qz;
            decode(cg, ci, qx, qy, qz);
            auto z1 = [&](uint32_t j) { return omega * (r[j] - m) * ldf(dinv[j]); };
            float rhs = 0.0f;
    #pragma unroll
            for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                for (int dy = 0; dy < 2; dy++)
    #pragma unroll
                    for (int dx = 0; dx < 2; dx++) {
                        const int x = 2 * qx + dx, y = 2 * qy + dy, z = 2 * qz + dz;
                        if (x >= fg.nx || y >= fg.ny || z >= fg.nz) continue;
                        const uint32_t i = (uint32_t)x + sy * (uint32_t)y + sz * (uint32_t)z;
                        const float di = ldf(dinv[i]);
                        if (di == 0.0f) continue;
                        const float rc_i = r[i] - m;
                        const float c = omega * rc_i * di;
                        uint32_t nb[6];
                        float w[6];
                        stencil6(fg, x, y, z, i, bx, by, bz, nb, w);
                        float a = 0.0f;
    #pragma unroll
                        for (int k = 0; k < 6; k++) a += w[k] * (c - z1(nb[k]));
                        rhs += rc_i - a;
                    }
            rc[ci] = cvt<VT>(rhs);
        }
    }
};

// ---------------------------------------------------------------- pass 4
// z2 = z1 + z_coarse[parent] (active), z = z2 + omega (r - m - A z2)/d; partial (r-m).z.
template <int V, class CT, class VT>
struct K_k_fused_smooth {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        Grid3 cg;
        const float* r;
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
        const VT* zc;
        VT* t;
        double* part;
        Scal* sc;
        int mode;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        Grid3 cg = A_.cg;
        const float* __restrict__ r = A_.r;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ bx = A_.bx;
        const CT* __restrict__ by = A_.by;
        const CT* __restrict__ bz = A_.bz;
        const VT* __restrict__ zc = A_.zc;
        VT* __restrict__ t = A_.t;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int mode = A_.mode;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
        const float m = (float)sc->mean;
        const float ps = (float)sc->pscale;  // coarse-correction scale (reference: 1)
        const uint32_t csy = (uint32_t)cg.nx, csz = (uint32_t)cg.nx * (uint32_t)cg.ny;
        uint32_t lo, hi;
        block_range(vg, g.n / V, lo, hi);
        double acc[1] = {0.0};
        // z2 at one cell (j at coordinates xq, yq, zq): 0 where inactive
        auto z2 = [&](uint32_t j, int xq, int yq, int zq) {
            const float dj = ldf(dinv[j]);
            if (dj == 0.0f) return 0.0f;
            return omega * (r[j] - m) * dj + ps * ldf(zc[(uint32_t)(xq >> 1) + csy * (uint32_t)(yq >> 1) + csz * (uint32_t)(zq >> 1)]);
        };
        for (uint32_t gi = lo + threadIdx.x; gi < hi; gi += MHDP_BLOCK) {
            const uint32_t i0 = gi * V;
            int x0, y, zz;
            decode(g, i0, x0, y, zz);
            const Nb n = neighbours<V>(g, x0, y, zz, i0);
            const W6<V, CT> W = weights<V, CT>(n, i0, bx, by, bz);
            // z2 along a row of V cells starting at j (row coordinates yq, zq)
            auto zrow = [&](uint32_t j, int yq, int zq) {
                const VF<V> rv = ldv<V>(r, j), dv = ldv<V>(dinv, j);
                const uint32_t cb = csy * (uint32_t)(yq >> 1) + csz * (uint32_t)(zq >> 1);
                VF<V> o;
    #pragma unroll
                for (int k = 0; k < V; k++)
                    o.v[k] = dv.v[k] != 0.0f ? omega * (rv.v[k] - m) * dv.v[k] + ps * ldf(zc[cb + (uint32_t)((x0 + k) >> 1)])
                                             : 0.0f;
                return o;
            };
            Star<V> s;
            {
                const VF<V> c = zrow(i0, y, zz), ym = zrow(n.ym, n.cym, zz), yp = zrow(n.yp, n.cyp, zz),
                            zm = zrow(n.zm, y, n.czm), zp = zrow(n.zp, y, n.czp);
                const float xm = z2(n.xm, n.cxm, y, zz), xp = z2(n.xp, n.cxp, y, zz);
    #pragma unroll
                for (int k = 0; k < V; k++) {
                    s.c[k] = c.v[k];
                    s.ym[k] = ym.v[k];
                    s.yp[k] = yp.v[k];
                    s.zm[k] = zm.v[k];
                    s.zp[k] = zp.v[k];
                    s.xm[k] = k > 0 ? c.v[k - 1] : xm;
                    s.xp[k] = k + 1 < V ? c.v[k + 1] : xp;
                }
            }
            const VF<V> rv = ldv<V>(r, i0), di = ldv<V>(dinv, i0);
            VF<V> tv;
    #pragma unroll
            for (int k = 0; k < V; k++) {
                float out = 0.0f;
                if (di.v[k] != 0.0f) {
                    const float rk = rv.v[k] - m;
                    out = s.c[k] + omega * (rk - star_action<V, CT>(W, s, k)) * di.v[k];
                    acc[0] += (double)rk * (double)out;
                }
                tv.v[k] = out;
            }
            stv<V>(t, i0, tv);
        }
        reduce_finish<1>(vg, acc, part, sc, 4, mode);
    }
};

struct K_k_clear_pending {
    static constexpr int MINB = 0;
    struct Args {
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid&) {
        if (threadIdx.x == 0) A_.sc->pending = 0u;
    }
};

// ---------------------------------------------------------------- launchers

void launch_clear_pending(hipStream_t s, Scal* sc) { go<K_k_clear_pending>(s, 1, {sc}); }

bool fused_vec4(const Grid3& g) { return (g.nx & 3) == 0; }

void launch_fused_apply(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                        const float* bx, const float* by, const float* bz, const float* dinv, float* q, double* part,
                        Scal* sc, int meanfree) {
    const bool v4 = fused_vec4(g);
    const dim3 B(MHDP_BLOCK), G(grid_blocks(v4 ? g.n / 4 : g.n));
    if (store == STORE_BF16) {
        if (v4) go<K_k_fused_apply<4, __hip_bfloat16>>(s, G.x, {g, (const __hip_bfloat16*)z, p0, p1, x, bx, by, bz, dinv, q, part, sc, meanfree});
        else go<K_k_fused_apply<1, __hip_bfloat16>>(s, G.x, {g, (const __hip_bfloat16*)z, p0, p1, x, bx, by, bz, dinv, q, part, sc, meanfree});
    } else {
        if (v4) go<K_k_fused_apply<4, float>>(s, G.x, {g, (const float*)z, p0, p1, x, bx, by, bz, dinv, q, part, sc, meanfree});
        else go<K_k_fused_apply<1, float>>(s, G.x, {g, (const float*)z, p0, p1, x, bx, by, bz, dinv, q, part, sc, meanfree});
    }
}
void launch_fused_xflush(hipStream_t s, const Grid3& g, const float* p0, const float* p1, float* x, const float* dinv,
                         Scal* sc, int meanfree) {
    const bool v4 = fused_vec4(g);
    const dim3 B(MHDP_BLOCK), G(grid_blocks(v4 ? g.n / 4 : g.n));
    if (v4) go<K_k_fused_xflush<4>>(s, G.x, {g, p0, p1, x, dinv, sc, meanfree});
    else go<K_k_fused_xflush<1>>(s, G.x, {g, p0, p1, x, dinv, sc, meanfree});
}
void launch_fused_update(hipStream_t s, const Grid3& g, float* r, const float* q, const float* dinv, double* part,
                         Scal* sc) {
    const bool v4 = fused_vec4(g);
    const dim3 B(MHDP_BLOCK), G(grid_blocks(v4 ? g.n / 4 : g.n));
    if (v4) go<K_k_fused_update<4>>(s, G.x, {g, r, q, dinv, part, sc});
    else go<K_k_fused_update<1>>(s, G.x, {g, r, q, dinv, part, sc});
}
void launch_fused_restrict(hipStream_t s, int store, const Grid3& cg, const Grid3& fg, const float* r, const void* dinv,
                           const void* bx, const void* by, const void* bz, void* rc, Scal* sc) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(cg.n));
    if (store == STORE_BF16)
        go<K_k_fused_restrict<__hip_bfloat16, __hip_bfloat16>>(s, G.x, {cg, fg, r, (const __hip_bfloat16*)dinv, (const __hip_bfloat16*)bx, (const __hip_bfloat16*)by, (const __hip_bfloat16*)bz, (__hip_bfloat16*)rc, sc});
    else
        go<K_k_fused_restrict<float, float>>(s, G.x, {cg, fg, r, (const float*)dinv, (const float*)bx, (const float*)by, (const float*)bz, (float*)rc, sc});
}
void launch_fused_smooth(hipStream_t s, int store, const Grid3& g, const Grid3& cg, const float* r, const void* dinv,
                         const void* bx, const void* by, const void* bz, const void* zc, void* t, double* part,
                         Scal* sc, int mode) {
    const bool v4 = fused_vec4(g);
    const dim3 B(MHDP_BLOCK), G(grid_blocks(v4 ? g.n / 4 : g.n));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        if (v4) go<K_k_fused_smooth<4, T, T>>(s, G.x, {g, cg, r, (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, (const T*)zc, (T*)t, part, sc, mode});
        else go<K_k_fused_smooth<1, T, T>>(s, G.x, {g, cg, r, (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, (const T*)zc, (T*)t, part, sc, mode});
    } else {
        using T = float;
        if (v4) go<K_k_fused_smooth<4, T, T>>(s, G.x, {g, cg, r, (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, (const T*)zc, (T*)t, part, sc, mode});
        else go<K_k_fused_smooth<1, T, T>>(s, G.x, {g, cg, r, (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, (const T*)zc, (T*)t, part, sc, mode});
    }
}
