// Brick variants of the fused fine-grid stencil passes: each thread owns a 4 x 2 x 2 brick
// of cells (four along x, two rows, two planes; needs nx % 4 == 0 and even ny, nz).  A
// brick is exactly two 2x2x2 aggregates, so the restriction needs no cross-thread sum, the
// prolongation reads two coarse values for the whole core, and twelve of the brick's
// eighteen y/z neighbour rows are the brick's own rows held in registers.  Loads per cell
// drop from ~11 (one group of 4 cells) to ~4.  Arithmetic and summation order per cell are
// those of fused.hip.
#include "mhdp_kernels.h"

#define CONVERGED(sc) ((sc)->conv > 0.5)

namespace {

struct F4 {
    float v[4];
};
__device__ __forceinline__ float bfl(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bfh(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }
__device__ __forceinline__ F4 ld4(const float* __restrict__ p, uint32_t i) {
    const float4 a = *reinterpret_cast<const float4*>(p + i);
    return F4{{a.x, a.y, a.z, a.w}};
}
__device__ __forceinline__ F4 ld4(const __hip_bfloat16* __restrict__ p, uint32_t i) {
    const uint2 a = *reinterpret_cast<const uint2*>(p + i);
    return F4{{bfl(a.x), bfh(a.x), bfl(a.y), bfh(a.y)}};
}
__device__ __forceinline__ void ld2(const float* __restrict__ p, uint32_t i, float& a, float& b) {
    const float2 v = *reinterpret_cast<const float2*>(p + i);
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void ld2(const __hip_bfloat16* __restrict__ p, uint32_t i, float& a, float& b) {
    const uint32_t v = *reinterpret_cast<const uint32_t*>(p + i);
    a = bfl(v);
    b = bfh(v);
}
__device__ __forceinline__ void st4(float* __restrict__ p, uint32_t i, const F4& a) {
    *reinterpret_cast<float4*>(p + i) = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]);
}
__device__ __forceinline__ uint32_t pk(float a, float b) {
    return (uint32_t)__bfloat16_as_ushort(__float2bfloat16(a)) |
           ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(b)) << 16);
}
__device__ __forceinline__ void st4(__hip_bfloat16* __restrict__ p, uint32_t i, const F4& a) {
    *reinterpret_cast<uint2*>(p + i) = make_uint2(pk(a.v[0], a.v[1]), pk(a.v[2], a.v[3]));
}
__device__ __forceinline__ void st2(float* __restrict__ p, uint32_t i, float a, float b) {
    *reinterpret_cast<float2*>(p + i) = make_float2(a, b);
}
__device__ __forceinline__ void st2(__hip_bfloat16* __restrict__ p, uint32_t i, float a, float b) {
    *reinterpret_cast<uint32_t*>(p + i) = pk(a, b);
}

// Geometry of the brick at (x0, y0, z0); every index stays in range (closed boundaries
// point at the brick itself and get weight 0).
struct Geo {
    int x0, y0, z0;
    uint32_t row[2][2];       // [dz][dy] start of the core rows
    uint32_t ymr[2], ypr[2];  // [dz] rows y0-1 and y0+2
    uint32_t zmr[2], zpr[2];  // [dy] rows in planes z0-1 and z0+2
    int ymy, ypy, zmz, zpz;
    int xm_off, xp_off, xmx, xpx;  // x-halo cells relative to x0
    bool xp_ok, yp_ok, zp_ok;
};
__device__ __forceinline__ Geo geo(const Grid3& g, uint32_t b) {
    Geo o;
    const uint32_t bx = (uint32_t)g.nx >> 2, by = (uint32_t)g.ny >> 1;
    const uint32_t gz = b / (bx * by), rem = b - gz * bx * by, gy = rem / bx, gx = rem - gy * bx;
    o.x0 = (int)(gx << 2);
    o.y0 = (int)(gy << 1);
    o.z0 = (int)(gz << 1);
    const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
    const uint32_t i0 = (uint32_t)o.x0 + sy * (uint32_t)o.y0 + sz * (uint32_t)o.z0;
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) o.row[dz][dy] = i0 + (uint32_t)dy * sy + (uint32_t)dz * sz;
    // y halo rows
    if (o.y0 > 0) { o.ymy = o.y0 - 1; for (int dz = 0; dz < 2; dz++) o.ymr[dz] = o.row[dz][0] - sy; }
    else if (g.per & 2u) { o.ymy = g.ny - 1; for (int dz = 0; dz < 2; dz++) o.ymr[dz] = o.row[dz][0] + (uint32_t)(g.ny - 1) * sy; }
    else { o.ymy = o.y0; for (int dz = 0; dz < 2; dz++) o.ymr[dz] = o.row[dz][0]; }
    if (o.y0 + 2 < g.ny) { o.ypy = o.y0 + 2; o.yp_ok = true; for (int dz = 0; dz < 2; dz++) o.ypr[dz] = o.row[dz][1] + sy; }
    else if (g.per & 2u) { o.ypy = 0; o.yp_ok = true; for (int dz = 0; dz < 2; dz++) o.ypr[dz] = o.row[dz][1] - (uint32_t)(g.ny - 1) * sy; }
    else { o.ypy = o.y0 + 1; o.yp_ok = false; for (int dz = 0; dz < 2; dz++) o.ypr[dz] = o.row[dz][1]; }
    // z halo rows
    if (o.z0 > 0) { o.zmz = o.z0 - 1; for (int dy = 0; dy < 2; dy++) o.zmr[dy] = o.row[0][dy] - sz; }
    else if (g.per & 4u) { o.zmz = g.nz - 1; for (int dy = 0; dy < 2; dy++) o.zmr[dy] = o.row[0][dy] + (uint32_t)(g.nz - 1) * sz; }
    else { o.zmz = o.z0; for (int dy = 0; dy < 2; dy++) o.zmr[dy] = o.row[0][dy]; }
    if (o.z0 + 2 < g.nz) { o.zpz = o.z0 + 2; o.zp_ok = true; for (int dy = 0; dy < 2; dy++) o.zpr[dy] = o.row[1][dy] + sz; }
    else if (g.per & 4u) { o.zpz = 0; o.zp_ok = true; for (int dy = 0; dy < 2; dy++) o.zpr[dy] = o.row[1][dy] - (uint32_t)(g.nz - 1) * sz; }
    else { o.zpz = o.z0 + 1; o.zp_ok = false; for (int dy = 0; dy < 2; dy++) o.zpr[dy] = o.row[1][dy]; }
    // x halo cells
    if (o.x0 > 0) { o.xmx = o.x0 - 1; }
    else if (g.per & 1u) { o.xmx = g.nx - 1; }
    else { o.xmx = o.x0; }
    if (o.x0 + 4 < g.nx) { o.xpx = o.x0 + 4; o.xp_ok = true; }
    else if (g.per & 1u) { o.xpx = 0; o.xp_ok = true; }
    else { o.xpx = o.x0 + 3; o.xp_ok = false; }
    o.xm_off = o.xmx - o.x0;
    o.xp_off = o.xpx - o.x0;
    return o;
}

// A field on the brick and its one-cell halo (faces only).
struct Field {
    float c[2][2][4];   // [dz][dy][k]
    float ym[2][4], yp[2][4];  // [dz][k]
    float zm[2][4], zp[2][4];  // [dy][k]
    float xm[2][2], xp[2][2];  // [dz][dy]
};
// rowf(start, y, z) -> F4 of the row starting at `start` (coordinates y, z of that row);
// ptf(index, x, y, z) -> value at one cell.
template <class RowF, class PtF>
__device__ __forceinline__ void gather(const Geo& o, Field& f, RowF&& rowf, PtF&& ptf) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const F4 a = rowf(o.row[dz][dy], o.y0 + dy, o.z0 + dz);
#pragma unroll
            for (int k = 0; k < 4; k++) f.c[dz][dy][k] = a.v[k];
            f.xm[dz][dy] = ptf(o.row[dz][dy] + (uint32_t)o.xm_off, o.xmx, o.y0 + dy, o.z0 + dz);
            f.xp[dz][dy] = ptf(o.row[dz][dy] + (uint32_t)o.xp_off, o.xpx, o.y0 + dy, o.z0 + dz);
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const F4 a = rowf(o.ymr[d], o.ymy, o.z0 + d), b = rowf(o.ypr[d], o.ypy, o.z0 + d);
        const F4 c = rowf(o.zmr[d], o.y0 + d, o.zmz), e = rowf(o.zpr[d], o.y0 + d, o.zpz);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            f.ym[d][k] = a.v[k];
            f.yp[d][k] = b.v[k];
            f.zm[d][k] = c.v[k];
            f.zp[d][k] = e.v[k];
        }
    }
}

// Face weights (lower faces of the core + the upper faces at the brick's high sides).
struct Wts {
    float bx[2][2][4], by[2][2][4], bz[2][2][4];
    float bxp[2][2];       // x+ face of the last cell of each row
    float byp[2][4];       // [dz] upper y faces of row dy=1
    float bzp[2][4];       // [dy] upper z faces of plane dz=1
};
template <class CT>
__device__ __forceinline__ void weights(const Geo& o, Wts& w, const CT* __restrict__ bx, const CT* __restrict__ by,
                                        const CT* __restrict__ bz) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const uint32_t r = o.row[dz][dy];
            const F4 a = ld4(bx, r), b = ld4(by, r), c = ld4(bz, r);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                w.bx[dz][dy][k] = a.v[k];
                w.by[dz][dy][k] = b.v[k];
                w.bz[dz][dy][k] = c.v[k];
            }
            w.bxp[dz][dy] = o.xp_ok ? ldf(bx[r + (uint32_t)o.xp_off]) : 0.0f;
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        F4 a{{0.f, 0.f, 0.f, 0.f}}, b{{0.f, 0.f, 0.f, 0.f}};
        if (o.yp_ok) a = ld4(by, o.ypr[d]);
        if (o.zp_ok) b = ld4(bz, o.zpr[d]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            w.byp[d][k] = a.v[k];
            w.bzp[d][k] = b.v[k];
        }
    }
}

// A f at core cell (dz, dy, k) in the reference order x-, x+, y-, y+, z-, z+.
__device__ __forceinline__ float act_at(const Wts& w, const Field& f, int dz, int dy, int k) {
    const float c = f.c[dz][dy][k];
    const float xm = k > 0 ? f.c[dz][dy][k - 1] : f.xm[dz][dy];
    const float xp = k < 3 ? f.c[dz][dy][k + 1] : f.xp[dz][dy];
    const float ym = dy == 1 ? f.c[dz][0][k] : f.ym[dz][k];
    const float yp = dy == 0 ? f.c[dz][1][k] : f.yp[dz][k];
    const float zm = dz == 1 ? f.c[0][dy][k] : f.zm[dy][k];
    const float zp = dz == 0 ? f.c[1][dy][k] : f.zp[dy][k];
    const float wxp = k < 3 ? w.bx[dz][dy][k + 1] : w.bxp[dz][dy];
    const float wyp = dy == 0 ? w.by[dz][1][k] : w.byp[dz][k];
    const float wzp = dz == 0 ? w.bz[1][dy][k] : w.bzp[dy][k];
    float a = 0.0f;
    a += w.bx[dz][dy][k] * (c - xm);
    a += wxp * (c - xp);
    a += w.by[dz][dy][k] * (c - ym);
    a += wyp * (c - yp);
    a += w.bz[dz][dy][k] * (c - zm);
    a += wzp * (c - zp);
    return a;
}

}  // namespace

bool brick_ok(const Grid3& g) { return (g.nx & 3) == 0 && (g.ny & 1) == 0 && (g.nz & 1) == 0; }
// MHDP_NO_BRICK=1 keeps the one-cell / four-cell kernels (comparison runs).
bool brick_enabled_env() {
    static const bool on = [] {
        const char* e = std::getenv("MHDP_NO_BRICK");
        return !(e && e[0] == '1');
    }();
    return on;
}

// pass 1 (k_fused_apply on bricks)
template <class VT, int MINB>
__global__ void __launch_bounds__(MHDP_BLOCK, MINB) k_brick_apply(Grid3 g, const VT* __restrict__ z, float* p0, float* p1,
                                                            float* __restrict__ x, const float* __restrict__ bx,
                                                            const float* __restrict__ by, const float* __restrict__ bz,
                                                            const float* __restrict__ dinv, float* __restrict__ q,
                                                            double* __restrict__ part, Scal* sc, int meanfree) {
    if (CONVERGED(sc)) return;
    const uint32_t it = (uint32_t)sc->iters;
    const float* __restrict__ pold = (it & 1u) ? p0 : p1;
    float* __restrict__ pnew = (it & 1u) ? p1 : p0;
    const float beta = (float)sc->beta;
    const bool pending = sc->pending != 0;
    const float alpha = (float)sc->alpha;
    const float pa = (pending && meanfree) ? pold[sc->anchor] : 0.0f;
    uint32_t lo, hi;
    block_range(g.n >> 4, lo, hi);
    double acc[1] = {0.0};
    for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
        const Geo o = geo(g, b);
        Wts w;
        weights(o, w, bx, by, bz);
        Field f;
        gather(
            o, f,
            [&](uint32_t j, int, int) {
                const F4 a = ld4(z, j), c = ld4(pold, j);
                return F4{{a.v[0] + beta * c.v[0], a.v[1] + beta * c.v[1], a.v[2] + beta * c.v[2], a.v[3] + beta * c.v[3]}};
            },
            [&](uint32_t j, int, int, int) { return ldf(z[j]) + beta * pold[j]; });
#pragma unroll
        for (int dz = 0; dz < 2; dz++)
#pragma unroll
            for (int dy = 0; dy < 2; dy++) {
                const uint32_t r = o.row[dz][dy];
                const F4 di = ld4(dinv, r);
                F4 qv, pn;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    pn.v[k] = di.v[k] != 0.0f ? f.c[dz][dy][k] : 0.0f;
                    qv.v[k] = act_at(w, f, dz, dy, k);
                    acc[0] += (double)pn.v[k] * (double)qv.v[k];
                }
                st4(pnew, r, pn);
                st4(q, r, qv);
                if (pending) {
                    const F4 po = ld4(pold, r);
                    F4 xv = ld4(x, r);
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (di.v[k] != 0.0f) xv.v[k] = xv.v[k] + alpha * (po.v[k] - pa);
                    st4(x, r, xv);
                }
            }
    }
    reduce_finish<1>(acc, part, sc, 2, FIN_ALPHA);
}

// pass 3 (k_fused_restrict on bricks): the two aggregates of a brick.  COARSE: a coarse level
// (Galerkin graph with reaction term; every child contributes, multigrid.wgsl:75-83; no mean).
template <class CT, class VT, class RT, bool COARSE, int MINB>
__global__ void __launch_bounds__(MHDP_BLOCK, MINB) k_brick_restrict(Grid3 cg, Grid3 g, const RT* __restrict__ r,
                                                                     const CT* __restrict__ dinv,
                                                                     const CT* __restrict__ bx,
                                                                     const CT* __restrict__ by,
                                                                     const CT* __restrict__ bz,
                                                                     const CT* __restrict__ react,
                                                                     VT* __restrict__ rc, Scal* sc) {
    if (CONVERGED(sc)) return;
    const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
    const float m = COARSE ? 0.0f : (float)sc->mean;
    uint32_t lo, hi;
    block_range(g.n >> 4, lo, hi);
    for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
        const Geo o = geo(g, b);
        Wts w;
        weights(o, w, bx, by, bz);
        Field f;
        gather(
            o, f,
            [&](uint32_t j, int, int) {
                const F4 a = ld4(r, j), d = ld4(dinv, j);
                F4 o4;
#pragma unroll
                for (int k = 0; k < 4; k++) o4.v[k] = omega * (a.v[k] - m) * d.v[k];
                return o4;
            },
            [&](uint32_t j, int, int, int) { return omega * (ldf(r[j]) - m) * ldf(dinv[j]); });
        float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
        for (int dz = 0; dz < 2; dz++)
#pragma unroll
            for (int dy = 0; dy < 2; dy++) {
                const uint32_t row = o.row[dz][dy];
                const F4 rv = ld4(r, row), di = ld4(dinv, row);
                F4 re{{0.f, 0.f, 0.f, 0.f}};
                if (COARSE) re = ld4(react, row);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (!COARSE && di.v[k] == 0.0f) continue;
                    float a = act_at(w, f, dz, dy, k);
                    if (COARSE) a = re.v[k] * f.c[dz][dy][k] + a;
                    const float v = (rv.v[k] - m) - a;
                    if (k < 2) s0 += v;
                    else s1 += v;
                }
            }
        const uint32_t ci = (uint32_t)(o.x0 >> 1) + (uint32_t)cg.nx * ((uint32_t)(o.y0 >> 1) + (uint32_t)cg.ny * (uint32_t)(o.z0 >> 1));
        st2(rc, ci, s0, s1);
    }
}

// pass 4 (k_fused_smooth on bricks).  COARSE: coarse level (reaction term, no mean, no r.z).
template <class CT, class VT, class RT, bool COARSE, int MINB>
__global__ void __launch_bounds__(MHDP_BLOCK, MINB) k_brick_smooth(Grid3 g, Grid3 cg, const RT* __restrict__ r,
                                                                   const CT* __restrict__ dinv,
                                                                   const CT* __restrict__ bx,
                                                                   const CT* __restrict__ by,
                                                                   const CT* __restrict__ bz,
                                                                   const CT* __restrict__ react,
                                                                   const VT* __restrict__ zc, VT* __restrict__ t,
                                                                   double* __restrict__ part, Scal* sc, int mode) {
    if (CONVERGED(sc)) return;
    const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
    const float m = COARSE ? 0.0f : (float)sc->mean;
    const float ps = (float)sc->pscale;  // coarse-correction scale (reference: 1)
    const uint32_t csy = (uint32_t)cg.nx, csz = (uint32_t)cg.nx * (uint32_t)cg.ny;
    uint32_t lo, hi;
    block_range(g.n >> 4, lo, hi);
    double acc[1] = {0.0};
    for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
        const Geo o = geo(g, b);
        Wts w;
        weights(o, w, bx, by, bz);
        const uint32_t cx = (uint32_t)(o.x0 >> 1);
        Field f;
        gather(
            o, f,
            [&](uint32_t j, int yq, int zq) {
                const F4 a = ld4(r, j), d = ld4(dinv, j);
                float c0, c1;
                ld2(zc, cx + csy * (uint32_t)(yq >> 1) + csz * (uint32_t)(zq >> 1), c0, c1);
                F4 o4;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    o4.v[k] = d.v[k] != 0.0f ? omega * (a.v[k] - m) * d.v[k] + ps * (k < 2 ? c0 : c1) : 0.0f;
                return o4;
            },
            [&](uint32_t j, int xq, int yq, int zq) {
                const float dj = ldf(dinv[j]);
                const float v = omega * (ldf(r[j]) - m) * dj +
                                ps * ldf(zc[(uint32_t)(xq >> 1) + csy * (uint32_t)(yq >> 1) + csz * (uint32_t)(zq >> 1)]);
                return dj != 0.0f ? v : 0.0f;
            });
#pragma unroll
        for (int dz = 0; dz < 2; dz++)
#pragma unroll
            for (int dy = 0; dy < 2; dy++) {
                const uint32_t row = o.row[dz][dy];
                const F4 rv = ld4(r, row), di = ld4(dinv, row);
                F4 re{{0.f, 0.f, 0.f, 0.f}};
                if (COARSE) re = ld4(react, row);
                F4 tv;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float out = 0.0f;
                    if (di.v[k] != 0.0f) {
                        const float rk = rv.v[k] - m;
                        float a = act_at(w, f, dz, dy, k);
                        if (COARSE) a = re.v[k] * f.c[dz][dy][k] + a;
                        out = f.c[dz][dy][k] + omega * (rk - a) * di.v[k];
                        if (!COARSE) acc[0] += (double)rk * (double)out;
                    }
                    tv.v[k] = out;
                }
                st4(t, row, tv);
            }
    }
    if constexpr (!COARSE) reduce_finish<1>(acc, part, sc, 4, mode);
}

// ---------------------------------------------------------------- launchers
// Minimum resident blocks per CU the register allocation targets (MHDP_BRICK_OCC: 2, 3, 4).
static int brick_occ() {
    static int occ = [] {
        const char* e = std::getenv("MHDP_BRICK_OCC");
        int v = e ? std::atoi(e) : 0;
        return (v == 2 || v == 3 || v == 4) ? v : 2;
    }();
    return occ;
}
#define BRICK_DISPATCH(KERNEL, TARGS, ...)                                                          \
    do {                                                                                          \
        switch (brick_occ()) {                                                                    \
        case 4: hipLaunchKernelGGL((KERNEL<TARGS, 4>), G, B, 0, s, __VA_ARGS__); break;             \
        case 3: hipLaunchKernelGGL((KERNEL<TARGS, 3>), G, B, 0, s, __VA_ARGS__); break;             \
        default: hipLaunchKernelGGL((KERNEL<TARGS, 2>), G, B, 0, s, __VA_ARGS__); break;            \
        }                                                                                         \
    } while (0)

void launch_brick_apply(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                        const float* bx, const float* by, const float* bz, const float* dinv, float* q, double* part,
                        Scal* sc, int meanfree) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16)
        BRICK_DISPATCH(k_brick_apply, __hip_bfloat16, g, (const __hip_bfloat16*)z, p0, p1, x, bx, by, bz, dinv, q, part,
                       sc, meanfree);
    else
        BRICK_DISPATCH(k_brick_apply, float, g, (const float*)z, p0, p1, x, bx, by, bz, dinv, q, part, sc, meanfree);
}
#define COMMA ,
void launch_brick_restrict(hipStream_t s, int store, const Grid3& cg, const Grid3& g, const float* r, const void* dinv,
                           const void* bx, const void* by, const void* bz, void* rc, Scal* sc) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_restrict, T COMMA T COMMA float COMMA false, cg, g, r, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)nullptr, (T*)rc, sc);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_restrict, T COMMA T COMMA float COMMA false, cg, g, r, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)nullptr, (T*)rc, sc);
    }
}
void launch_brick_smooth(hipStream_t s, int store, const Grid3& g, const Grid3& cg, const float* r, const void* dinv,
                         const void* bx, const void* by, const void* bz, const void* zc, void* t, double* part,
                         Scal* sc, int mode) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA float COMMA false, g, cg, r, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)nullptr, (const T*)zc, (T*)t, part, sc, mode);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA float COMMA false, g, cg, r, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)nullptr, (const T*)zc, (T*)t, part, sc, mode);
    }
}
// Coarse level l (Galerkin graph, residual in the cycle's storage type).
void launch_brick_restrict_coarse(hipStream_t s, int store, const Grid3& cg, const Grid3& g, const void* r,
                                  const void* dinv, const void* bx, const void* by, const void* bz, const void* react,
                                  void* rc, Scal* sc) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_restrict, T COMMA T COMMA T COMMA true, cg, g, (const T*)r, (const T*)dinv,
                       (const T*)bx, (const T*)by, (const T*)bz, (const T*)react, (T*)rc, sc);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_restrict, T COMMA T COMMA T COMMA true, cg, g, (const T*)r, (const T*)dinv,
                       (const T*)bx, (const T*)by, (const T*)bz, (const T*)react, (T*)rc, sc);
    }
}
void launch_brick_smooth_coarse(hipStream_t s, int store, const Grid3& g, const Grid3& cg, const void* r,
                                const void* dinv, const void* bx, const void* by, const void* bz, const void* react,
                                const void* zc, void* t, Scal* sc) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA T COMMA true, g, cg, (const T*)r, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)react, (const T*)zc, (T*)t, (double*)nullptr, sc, 0);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA T COMMA true, g, cg, (const T*)r, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)react, (const T*)zc, (T*)t, (double*)nullptr, sc, 0);
    }
}
