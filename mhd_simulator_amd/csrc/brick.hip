// Brick variants of the fused fine-grid stencil passes: each thread owns a 4 x 2 x 2 brick
// of cells (four along x, two rows, two planes; needs nx % 4 == 0 and even ny, nz).  A
// brick is exactly two 2x2x2 aggregates, so the restriction needs no cross-thread sum, the
// prolongation reads two coarse values for the whole core, and twelve of the brick's
// eighteen y/z neighbour rows are the brick's own rows held in registers.  Loads per cell
// drop from ~11 (one group of 4 cells) to ~4.  Arithmetic and summation order per cell are
// those of fused.hip.
#include "mhdp_kernels.h"
#include "launch.h"
#include "raw.h"
#include "mob.h"

#define CONVERGED(sc) ((sc)->conv > 0.5)

namespace {

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
// One array on the brick and its face halo: 4 core rows, 8 halo rows, 8 halo cells.
template <class T>
struct Rows {
    R4<T> c[2][2];            // [dz][dy]
    R4<T> ym[2], yp[2];       // [dz]
    R4<T> zm[2], zp[2];       // [dy]
    R1<T> xm[2][2], xp[2][2]; // [dz][dy]
};
template <class T>
__device__ __forceinline__ void load_rows(const Geo& o, const T* __restrict__ p, Rows<T>& r) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            r.c[dz][dy] = ldr4(p, o.row[dz][dy]);
            r.xm[dz][dy] = ldr1(p, o.row[dz][dy] + (uint32_t)o.xm_off);
            r.xp[dz][dy] = ldr1(p, o.row[dz][dy] + (uint32_t)o.xp_off);
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        r.ym[d] = ldr4(p, o.ymr[d]);
        r.yp[d] = ldr4(p, o.ypr[d]);
        r.zm[d] = ldr4(p, o.zmr[d]);
        r.zp[d] = ldr4(p, o.zpr[d]);
    }
}
// The rows of an array the pass must not read, as zeros.
__device__ __forceinline__ void zero_rows(Rows<float>& r) {
    const R4<float> z4{make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            r.c[dz][dy] = z4;
            r.xm[dz][dy] = R1<float>{0.0f};
            r.xp[dz][dy] = R1<float>{0.0f};
        }
#pragma unroll
    for (int d = 0; d < 2; d++) r.ym[d] = r.yp[d] = r.zm[d] = r.zp[d] = z4;
}
template <class T>
__device__ __forceinline__ void fence_rows(Rows<T>& r) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            fence(r.c[dz][dy]);
            fence(r.xm[dz][dy]);
            fence(r.xp[dz][dy]);
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        fence(r.ym[d]);
        fence(r.yp[d]);
        fence(r.zm[d]);
        fence(r.zp[d]);
    }
}
// Field from two arrays A, B: rows through fr(a, b, row-id) -> F4, cells through fp(a, b, id).
// Row ids: 0..3 core (dz*2+dy), 4+d ym, 6+d yp, 8+d zm, 10+d zp; cell ids 0..3 xm, 4..7 xp.
template <class TA, class TB, class FR, class FP>
__device__ __forceinline__ void field2(const Rows<TA>& a, const Rows<TB>& b, Field& f, FR&& fr, FP&& fp) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const F4 v = fr(un(a.c[dz][dy]), un(b.c[dz][dy]), dz * 2 + dy);
#pragma unroll
            for (int k = 0; k < 4; k++) f.c[dz][dy][k] = v.v[k];
            f.xm[dz][dy] = fp(un(a.xm[dz][dy]), un(b.xm[dz][dy]), dz * 2 + dy);
            f.xp[dz][dy] = fp(un(a.xp[dz][dy]), un(b.xp[dz][dy]), 4 + dz * 2 + dy);
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const F4 v0 = fr(un(a.ym[d]), un(b.ym[d]), 4 + d), v1 = fr(un(a.yp[d]), un(b.yp[d]), 6 + d);
        const F4 v2 = fr(un(a.zm[d]), un(b.zm[d]), 8 + d), v3 = fr(un(a.zp[d]), un(b.zp[d]), 10 + d);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            f.ym[d][k] = v0.v[k];
            f.yp[d][k] = v1.v[k];
            f.zm[d][k] = v2.v[k];
            f.zp[d][k] = v3.v[k];
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
struct WRaw {
    R4<CT> bx[2][2], by[2][2], bz[2][2];
    R1<CT> bxp[2][2];
    R4<CT> byp[2], bzp[2];
};
template <class CT>
__device__ __forceinline__ void load_weights(const Geo& o, WRaw<CT>& w, const CT* __restrict__ bx,
                                             const CT* __restrict__ by, const CT* __restrict__ bz) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const uint32_t r = o.row[dz][dy];
            w.bx[dz][dy] = ldr4(bx, r);
            w.by[dz][dy] = ldr4(by, r);
            w.bz[dz][dy] = ldr4(bz, r);
            w.bxp[dz][dy] = ldr1(bx, r + (uint32_t)o.xp_off);
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        w.byp[d] = ldr4(by, o.ypr[d]);
        w.bzp[d] = ldr4(bz, o.zpr[d]);
    }
}
template <class CT>
__device__ __forceinline__ void fence_weights(WRaw<CT>& w) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            fence(w.bx[dz][dy]);
            fence(w.by[dz][dy]);
            fence(w.bz[dz][dy]);
            fence(w.bxp[dz][dy]);
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        fence(w.byp[d]);
        fence(w.bzp[d]);
    }
}
// Closed high-side faces (non-periodic boundary) get weight 0.
template <class CT>
__device__ __forceinline__ void weights(const Geo& o, const WRaw<CT>& raw, Wts& w) {  // FP32 weights
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const F4 a = un(raw.bx[dz][dy]), b = un(raw.by[dz][dy]), c = un(raw.bz[dz][dy]);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                w.bx[dz][dy][k] = a.v[k];
                w.by[dz][dy][k] = b.v[k];
                w.bz[dz][dy][k] = c.v[k];
            }
            w.bxp[dz][dy] = o.xp_ok ? un(raw.bxp[dz][dy]) : 0.0f;
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        const F4 a = un(raw.byp[d]), b = un(raw.bzp[d]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            w.byp[d][k] = o.yp_ok ? a.v[k] : 0.0f;
            w.bzp[d][k] = o.zp_ok ? b.v[k] : 0.0f;
        }
    }
}

// BF16 mobilities stay packed two to a register until use (halves their register footprint).
struct WtsH {
    uint32_t bx[2][2][2], by[2][2][2], bz[2][2][2];  // [dz][dy][word], word w holds lanes 2w, 2w+1
    float bxp[2][2];
    uint32_t byp[2][2], bzp[2][2];                   // [dz or dy][word]
};
__device__ __forceinline__ void weights(const Geo& o, const WRaw<__hip_bfloat16>& raw, WtsH& w) {
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            w.bx[dz][dy][0] = raw.bx[dz][dy].v.x; w.bx[dz][dy][1] = raw.bx[dz][dy].v.y;
            w.by[dz][dy][0] = raw.by[dz][dy].v.x; w.by[dz][dy][1] = raw.by[dz][dy].v.y;
            w.bz[dz][dy][0] = raw.bz[dz][dy].v.x; w.bz[dz][dy][1] = raw.bz[dz][dy].v.y;
            w.bxp[dz][dy] = o.xp_ok ? un(raw.bxp[dz][dy]) : 0.0f;
        }
#pragma unroll
    for (int d = 0; d < 2; d++) {
        w.byp[d][0] = o.yp_ok ? raw.byp[d].v.x : 0u;
        w.byp[d][1] = o.yp_ok ? raw.byp[d].v.y : 0u;
        w.bzp[d][0] = o.zp_ok ? raw.bzp[d].v.x : 0u;
        w.bzp[d][1] = o.zp_ok ? raw.bzp[d].v.y : 0u;
    }
}
__device__ __forceinline__ float hx(const uint32_t (&w)[2], int k) { return (k & 1) ? bfh(w[k >> 1]) : bfl(w[k >> 1]); }
// the face weights of core cell (dz, dy, k): lower x, upper x, lower y, upper y, lower z, upper z
__device__ __forceinline__ void face_w(const Wts& w, int dz, int dy, int k, float (&o)[6]) {
    o[0] = w.bx[dz][dy][k];
    o[1] = k < 3 ? w.bx[dz][dy][k + 1] : w.bxp[dz][dy];
    o[2] = w.by[dz][dy][k];
    o[3] = dy == 0 ? w.by[dz][1][k] : w.byp[dz][k];
    o[4] = w.bz[dz][dy][k];
    o[5] = dz == 0 ? w.bz[1][dy][k] : w.bzp[dy][k];
}
__device__ __forceinline__ void face_w(const WtsH& w, int dz, int dy, int k, float (&o)[6]) {
    o[0] = hx(w.bx[dz][dy], k);
    o[1] = k < 3 ? hx(w.bx[dz][dy], k + 1) : w.bxp[dz][dy];
    o[2] = hx(w.by[dz][dy], k);
    o[3] = dy == 0 ? hx(w.by[dz][1], k) : hx(w.byp[dz], k);
    o[4] = hx(w.bz[dz][dy], k);
    o[5] = dz == 0 ? hx(w.bz[1][dy], k) : hx(w.bzp[dy], k);
}
template <class CT> struct WtsOf { using type = Wts; };
template <> struct WtsOf<__hip_bfloat16> { using type = WtsH; };

// A f at core cell (dz, dy, k) in the reference order x-, x+, y-, y+, z-, z+.
template <class W>
__device__ __forceinline__ float act_at(const W& w, const Field& f, int dz, int dy, int k) {
    const float c = f.c[dz][dy][k];
    const float xm = k > 0 ? f.c[dz][dy][k - 1] : f.xm[dz][dy];
    const float xp = k < 3 ? f.c[dz][dy][k + 1] : f.xp[dz][dy];
    const float ym = dy == 1 ? f.c[dz][0][k] : f.ym[dz][k];
    const float yp = dy == 0 ? f.c[dz][1][k] : f.yp[dz][k];
    const float zm = dz == 1 ? f.c[0][dy][k] : f.zm[dy][k];
    const float zp = dz == 0 ? f.c[1][dy][k] : f.zp[dy][k];
    float fw[6];
    face_w(w, dz, dy, k, fw);
    float a = 0.0f;
    a += fw[0] * (c - xm);
    a += fw[1] * (c - xp);
    a += fw[2] * (c - ym);
    a += fw[3] * (c - yp);
    a += fw[4] * (c - zm);
    a += fw[5] * (c - zp);
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

// pass 1 (k_fused_apply on bricks).  The reciprocal diagonal only marks the active cells here
// (tested against zero), so the mixed solver passes its BF16 copy (VT): 1/diag lies in
// [1e-5, 1e30], where BF16 rounding never gives zero.
template <class VT, int MB_>
struct K_k_brick_apply {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 g;
        const VT* z;
        float* p0;
        float* p1;
        float* x;
        const float* bx;
        const float* by;
        const float* bz;
        const VT* dinv;
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
        const VT* __restrict__ dinv = A_.dinv;
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
        // a solve's first iteration (beta = 0, nothing owed) takes the old search direction as
        // zeros: a refinement step's start leaves it unset (the loads stay unconditional, so that
        // they issue with the others and no wait depends on a branch)
        const bool first = it == 0u;
        uint32_t lo, hi;
        block_range(vg, g.n >> 4, lo, hi);
        double acc[2] = {0.0, 0.0};
        for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
            const Geo o = geo(g, b);
            Rows<VT> zr;
            Rows<float> pr;
            WRaw<float> wr;
            R4<VT> dr[2][2];
            R4<float> xr[2][2];
            load_rows(o, z, zr);
            load_rows(o, pold, pr);
            load_weights(o, wr, bx, by, bz);
    #pragma unroll
            for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    dr[dz][dy] = ldr4(dinv, o.row[dz][dy]);
                    xr[dz][dy] = ldr4(x, o.row[dz][dy]);
                }
            fence_rows(zr);
            fence_rows(pr);
            fence_weights(wr);
    #pragma unroll
            for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    fence(dr[dz][dy]);
                    fence(xr[dz][dy]);
                }
            if (first) zero_rows(pr);
            Wts w;
            weights(o, wr, w);
            Field f;
            field2(
                zr, pr, f,
                [&](const F4& a, const F4& c, int) {
                    return F4{{a.v[0] + beta * c.v[0], a.v[1] + beta * c.v[1], a.v[2] + beta * c.v[2], a.v[3] + beta * c.v[3]}};
                },
                [&](float a, float c, int) { return a + beta * c; });
    #pragma unroll
            for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const uint32_t r = o.row[dz][dy];
                    const F4 di = un(dr[dz][dy]);
                    F4 qv, pn;
    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        pn.v[k] = di.v[k] != 0.0f ? f.c[dz][dy][k] : 0.0f;
                        qv.v[k] = act_at(w, f, dz, dy, k);
                        acc[0] += (double)pn.v[k] * (double)qv.v[k];
                        acc[1] += di.v[k] != 0.0f ? (double)qv.v[k] : 0.0;
                    }
                    st4(pnew, r, pn);
                    st4(q, r, qv);
                    if (pending) {
                        const F4 po = un(pr.c[dz][dy]);
                        F4 xv = un(xr[dz][dy]);
    #pragma unroll
                        for (int k = 0; k < 4; k++)
                            if (di.v[k] != 0.0f) xv.v[k] = xv.v[k] + alpha * (po.v[k] - pa);
                        st4(x, r, xv);
                    }
                }
        }
        reduce_finish<2>(vg, acc, part, sc, 2, FIN_ALPHA);
    }
};

// pass 1 with the face mobilities formed in the pass instead of loaded: from the liquid fraction
// of each cell (two-phase) and the per-cell open-face bytes (bit a: the cell's lower face on axis a
// is open, bit 3: fluid, bit 4: a component's anchor).  The operator is the stored one bit for bit:
// density() and mobility() (solver.wgsl:40-51) in the coefficient pass's FP32 operations, rounded
// one by one (no fused multiply-add), and a cell is active where the coefficient pass gives it a
// nonzero reciprocal diagonal (fluid; in the pinned gauge not an anchor).  The pass reads 5 bytes
// per cell of coefficients instead of 14 (three FP32 mobilities and the BF16 1/diag).
template <class VT, bool TWO, int MB_>
struct K_k_brick_apply_m {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 g;
        const VT* z;
        float* p0;
        float* p1;
        float* x;
        const float* alpha;  // liquid fraction (TWO)
        const uint8_t* open8;
        MobP mp;
        float* q;
        double* part;
        Scal* sc;
        int meanfree;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        const Grid3 g = A_.g;
        const VT* __restrict__ z = A_.z;
        float* __restrict__ x = A_.x;
        const uint8_t* __restrict__ open8 = A_.open8;
        const MobP mp = A_.mp;
        float* __restrict__ q = A_.q;
        Scal* sc = A_.sc;
        const int meanfree = A_.meanfree;
        if (CONVERGED(sc)) return;
        const uint32_t it = (uint32_t)sc->iters;
        const float* __restrict__ pold = (it & 1u) ? A_.p0 : A_.p1;
        float* __restrict__ pnew = (it & 1u) ? A_.p1 : A_.p0;
        const float beta = (float)sc->beta;
        const bool pending = sc->pending != 0;
        const float alpha = (float)sc->alpha;
        const float pa = (pending && meanfree) ? pold[sc->anchor] : 0.0f;
        // a solve's first iteration (beta = 0, nothing owed) takes the old search direction as
        // zeros: a refinement step's start leaves it unset (the loads stay unconditional, so that
        // they issue with the others and no wait depends on a branch)
        const bool first = it == 0u;
        uint32_t lo, hi;
        block_range(vg, g.n >> 4, lo, hi);
        double acc[2] = {0.0, 0.0};
        for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
            const Geo o = geo(g, b);
            Rows<VT> zr;
            Rows<float> pr, ar;
            R4<float> xr[2][2];
            uint32_t ob[2][2], obx[2][2], oby[2], obz[2];
            load_rows(o, z, zr);
            load_rows(o, pold, pr);
            if (TWO) load_rows(o, A_.alpha, ar);
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const uint32_t r = o.row[dz][dy];
                    xr[dz][dy] = ldr4(x, r);
                    ob[dz][dy] = *reinterpret_cast<const uint32_t*>(open8 + r);
                    obx[dz][dy] = open8[r + (uint32_t)o.xp_off];
                }
#pragma unroll
            for (int d = 0; d < 2; d++) {
                oby[d] = *reinterpret_cast<const uint32_t*>(open8 + o.ypr[d]);
                obz[d] = *reinterpret_cast<const uint32_t*>(open8 + o.zpr[d]);
            }
            fence_rows(zr);
            fence_rows(pr);
            if (TWO) fence_rows(ar);
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    fence(xr[dz][dy]);
                    asm volatile("" : "+v"(ob[dz][dy]), "+v"(obx[dz][dy]));
                }
#pragma unroll
            for (int d = 0; d < 2; d++) asm volatile("" : "+v"(oby[d]), "+v"(obz[d]));
            if (first) zero_rows(pr);
            // densities of the brick and its face halo, then the mobilities of every face
            Field dn;
            if (TWO) field2(ar, ar, dn, [&](const F4& a, const F4&, int) {
                    return F4{{mob_dens(mp, a.v[0]), mob_dens(mp, a.v[1]), mob_dens(mp, a.v[2]), mob_dens(mp, a.v[3])}};
                }, [&](float a, float, int) { return mob_dens(mp, a); });
            Wts w;
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const uint32_t bits = obyte(ob[dz][dy], k);
                        const float dc = TWO ? dn.c[dz][dy][k] : 0.0f;
                        const float dxm = TWO ? (k > 0 ? dn.c[dz][dy][k - 1] : dn.xm[dz][dy]) : 0.0f;
                        const float dym = TWO ? (dy == 1 ? dn.c[dz][0][k] : dn.ym[dz][k]) : 0.0f;
                        const float dzm = TWO ? (dz == 1 ? dn.c[0][dy][k] : dn.zm[dy][k]) : 0.0f;
                        w.bx[dz][dy][k] = mob_face<TWO>(mp, bits & 1u, dxm, dc);
                        w.by[dz][dy][k] = mob_face<TWO>(mp, bits & 2u, dym, dc);
                        w.bz[dz][dy][k] = mob_face<TWO>(mp, bits & 4u, dzm, dc);
                    }
                    w.bxp[dz][dy] = mob_face<TWO>(mp, o.xp_ok && (obx[dz][dy] & 1u), TWO ? dn.c[dz][dy][3] : 0.0f,
                                                  TWO ? dn.xp[dz][dy] : 0.0f);
                }
#pragma unroll
            for (int d = 0; d < 2; d++)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    w.byp[d][k] = mob_face<TWO>(mp, o.yp_ok && (obyte(oby[d], k) & 2u), TWO ? dn.c[d][1][k] : 0.0f,
                                                TWO ? dn.yp[d][k] : 0.0f);
                    w.bzp[d][k] = mob_face<TWO>(mp, o.zp_ok && (obyte(obz[d], k) & 4u), TWO ? dn.c[1][d][k] : 0.0f,
                                                TWO ? dn.zp[d][k] : 0.0f);
                }
            Field f;
            field2(
                zr, pr, f,
                [&](const F4& a, const F4& c, int) {
                    return F4{{a.v[0] + beta * c.v[0], a.v[1] + beta * c.v[1], a.v[2] + beta * c.v[2], a.v[3] + beta * c.v[3]}};
                },
                [&](float a, float c, int) { return a + beta * c; });
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const uint32_t r = o.row[dz][dy];
                    F4 qv, pn;
                    bool act[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const uint32_t bits = obyte(ob[dz][dy], k);
                        act[k] = open_active(bits, meanfree);
                        pn.v[k] = act[k] ? f.c[dz][dy][k] : 0.0f;
                        qv.v[k] = act_at(w, f, dz, dy, k);
                        acc[0] += (double)pn.v[k] * (double)qv.v[k];
                        acc[1] += act[k] ? (double)qv.v[k] : 0.0;
                    }
                    st4(pnew, r, pn);
                    st4(q, r, qv);
                    if (pending) {
                        const F4 po = un(pr.c[dz][dy]);
                        F4 xv = un(xr[dz][dy]);
#pragma unroll
                        for (int k = 0; k < 4; k++)
                            if (act[k]) xv.v[k] = xv.v[k] + alpha * (po.v[k] - pa);
                        st4(x, r, xv);
                    }
                }
        }
        reduce_finish<2>(vg, acc, A_.part, sc, 2, FIN_ALPHA);
    }
};

// pass 3 (k_fused_restrict on bricks): the two aggregates of a brick.  COARSE: a coarse level
// (Galerkin graph with reaction term; every child contributes, multigrid.wgsl:75-83; no mean).
template <class CT, class VT, class RT, bool COARSE, int MB_>
struct K_k_brick_restrict {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 cg;
        Grid3 g;
        const RT* r;
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
        const CT* react;
        VT* rc;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 cg = A_.cg;
        Grid3 g = A_.g;
        const RT* __restrict__ r = A_.r;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ bx = A_.bx;
        const CT* __restrict__ by = A_.by;
        const CT* __restrict__ bz = A_.bz;
        const CT* __restrict__ react = A_.react;
        VT* __restrict__ rc = A_.rc;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
        const float m = COARSE ? 0.0f : (float)sc->mean;
        uint32_t lo, hi;
        block_range(vg, g.n >> 4, lo, hi);
        for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
            const Geo o = geo(g, b);
            Rows<RT> rr;
            Rows<CT> dd;
            WRaw<CT> wr;
            R4<CT> er[2][2];
            load_rows(o, r, rr);
            load_rows(o, dinv, dd);
            load_weights(o, wr, bx, by, bz);
            if (COARSE) {
    #pragma unroll
                for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                    for (int dy = 0; dy < 2; dy++) er[dz][dy] = ldr4(react, o.row[dz][dy]);
            }
            fence_rows(rr);
            fence_rows(dd);
            fence_weights(wr);
            if (COARSE) {
    #pragma unroll
                for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                    for (int dy = 0; dy < 2; dy++) fence(er[dz][dy]);
            }
            typename WtsOf<CT>::type w;
            weights(o, wr, w);
            Field f;
            field2(
                rr, dd, f,
                [&](const F4& a, const F4& d, int) {
                    F4 o4;
    #pragma unroll
                    for (int k = 0; k < 4; k++) o4.v[k] = omega * (a.v[k] - m) * d.v[k];
                    return o4;
                },
                [&](float a, float d, int) { return omega * (a - m) * d; });
            float s0 = 0.0f, s1 = 0.0f;
    #pragma unroll
            for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const F4 rv = un(rr.c[dz][dy]), di = un(dd.c[dz][dy]);
                    F4 re{{0.f, 0.f, 0.f, 0.f}};
                    if (COARSE) re = un(er[dz][dy]);
    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        float a = act_at(w, f, dz, dy, k);
                        if (COARSE) a = re.v[k] * f.c[dz][dy][k] + a;
                        float v = (rv.v[k] - m) - a;
                        if (!COARSE) v = di.v[k] != 0.0f ? v : 0.0f;
                        if (k < 2) s0 += v;
                        else s1 += v;
                    }
                }
            const uint32_t ci = (uint32_t)(o.x0 >> 1) + (uint32_t)cg.nx * ((uint32_t)(o.y0 >> 1) + (uint32_t)cg.ny * (uint32_t)(o.z0 >> 1));
            st2(rc, ci, s0, s1);
        }
    }
};

// passes 2+3 in one: the residual update r' = (r - mean) - alpha q (k_fused_update's
// arithmetic) evaluated on the brick and its halo, r' stored for the brick's own cells in the
// twin buffer, and the restriction of the pre-smoothed residual of r' - mean'; mean' is the
// mean of r' known from the apply pass.  Partials (0, sum r'^2, 0, sum r') -> FIN_RESTRICT_UPD.
template <class CT, class VT, int MB_>
struct K_k_brick_restrict_update {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 cg;
        Grid3 g;
        float* r0;
        float* r1;
        const float* q;
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
        VT* rc;
        double* part;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 cg = A_.cg;
        Grid3 g = A_.g;
        float* r0 = A_.r0;
        float* r1 = A_.r1;
        const float* __restrict__ q = A_.q;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ bx = A_.bx;
        const CT* __restrict__ by = A_.by;
        const CT* __restrict__ bz = A_.bz;
        VT* __restrict__ rc = A_.rc;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;
        const float alpha = (float)sc->alpha;
        const float m = (float)sc->mean;        // mean of r (deferred projection)
        const float mn = (float)sc->mean_next;  // mean of r'
        const float* __restrict__ rin = sc->rsel ? r1 : r0;
        float* __restrict__ rout = sc->rsel ? r0 : r1;
        uint32_t lo, hi;
        block_range(vg, g.n >> 4, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
            const Geo o = geo(g, b);
            Rows<float> rr, qq;
            Rows<CT> dd;
            WRaw<CT> wr;
            load_rows(o, rin, rr);
            load_rows(o, q, qq);
            load_rows(o, dinv, dd);
            load_weights(o, wr, bx, by, bz);
            fence_rows(rr);
            fence_rows(qq);
            fence_rows(dd);
            fence_weights(wr);
            typename WtsOf<CT>::type w;
            weights(o, wr, w);
            // r' on the brick and its halo; the pre-smoothed z1 = omega (r' - mean') / diag
            Field rn, f;
            field2(
                rr, qq, rn,
                [&](const F4& a, const F4& c, int) {
                    F4 o4;
    #pragma unroll
                    for (int k = 0; k < 4; k++) o4.v[k] = (a.v[k] - m) - alpha * c.v[k];
                    return o4;
                },
                [&](float a, float c, int) { return (a - m) - alpha * c; });
            {
                Rows<CT>& d = dd;
    #pragma unroll
                for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                    for (int dy = 0; dy < 2; dy++) {
                        const F4 dv = un(d.c[dz][dy]);
    #pragma unroll
                        for (int k = 0; k < 4; k++) {
                            rn.c[dz][dy][k] = dv.v[k] != 0.0f ? rn.c[dz][dy][k] : 0.0f;
                            f.c[dz][dy][k] = omega * (rn.c[dz][dy][k] - mn) * dv.v[k];
                        }
                        const float dxm = un(d.xm[dz][dy]), dxp = un(d.xp[dz][dy]);
                        f.xm[dz][dy] = omega * ((dxm != 0.0f ? rn.xm[dz][dy] : 0.0f) - mn) * dxm;
                        f.xp[dz][dy] = omega * ((dxp != 0.0f ? rn.xp[dz][dy] : 0.0f) - mn) * dxp;
                    }
    #pragma unroll
                for (int e = 0; e < 2; e++) {
                    const F4 a = un(d.ym[e]), b2 = un(d.yp[e]), c = un(d.zm[e]), h = un(d.zp[e]);
    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        f.ym[e][k] = omega * ((a.v[k] != 0.0f ? rn.ym[e][k] : 0.0f) - mn) * a.v[k];
                        f.yp[e][k] = omega * ((b2.v[k] != 0.0f ? rn.yp[e][k] : 0.0f) - mn) * b2.v[k];
                        f.zm[e][k] = omega * ((c.v[k] != 0.0f ? rn.zm[e][k] : 0.0f) - mn) * c.v[k];
                        f.zp[e][k] = omega * ((h.v[k] != 0.0f ? rn.zp[e][k] : 0.0f) - mn) * h.v[k];
                    }
                }
            }
            float s0 = 0.0f, s1 = 0.0f;
    #pragma unroll
            for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const F4 di = un(dd.c[dz][dy]);
                    F4 ro;
    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float rk = rn.c[dz][dy][k];
                        ro.v[k] = rk;
                        acc[1] += (double)rk * (double)rk;
                        acc[3] += (double)rk;
                        const float a = act_at(w, f, dz, dy, k);
                        float v = (rk - mn) - a;
                        v = di.v[k] != 0.0f ? v : 0.0f;
                        if (k < 2) s0 += v;
                        else s1 += v;
                    }
                    st4(rout, o.row[dz][dy], ro);
                }
            const uint32_t ci = (uint32_t)(o.x0 >> 1) + (uint32_t)cg.nx * ((uint32_t)(o.y0 >> 1) + (uint32_t)cg.ny * (uint32_t)(o.z0 >> 1));
            st2(rc, ci, s0, s1);
        }
        reduce_finish<4>(vg, acc, part, sc, 3, FIN_RESTRICT_UPD);
    }
};

// pass 4 (k_fused_smooth on bricks).  COARSE: coarse level (reaction term, no mean, no r.z).
template <class CT, class VT, class RT, bool COARSE, int MB_>
struct K_k_brick_smooth {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 g;
        Grid3 cg;
        const RT* r0;
        const RT* r1;
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
        const CT* react;
        const VT* zc;
        VT* t;
        double* part;
        Scal* sc;
        int mode;
        int use_next;
        const VT* base;  // COARSE: when set, t = base + theta * (smoothed value); base may be t
        float theta;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        Grid3 cg = A_.cg;
        const RT* r0 = A_.r0;
        const RT* r1 = A_.r1;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ bx = A_.bx;
        const CT* __restrict__ by = A_.by;
        const CT* __restrict__ bz = A_.bz;
        const CT* __restrict__ react = A_.react;
        const VT* __restrict__ zc = A_.zc;
        VT* __restrict__ t = A_.t;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int mode = A_.mode;
        int use_next = A_.use_next;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;  // smoother weight (reference: 2/3)
        const float m = COARSE ? 0.0f : (float)(use_next ? sc->mean_next : sc->mean);
        const float ps = (float)sc->pscale;  // coarse-correction scale (reference: 1)
        const RT* __restrict__ r = (!COARSE && sc->rsel) ? r1 : r0;
        const uint32_t csy = (uint32_t)cg.nx, csz = (uint32_t)cg.nx * (uint32_t)cg.ny;
        uint32_t lo, hi;
        block_range(vg, g.n >> 4, lo, hi);
        double acc[1] = {0.0};
        for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
            const Geo o = geo(g, b);
            Rows<RT> rr;
            Rows<CT> dd;
            WRaw<CT> wr;
            R4<CT> er[2][2];
            load_rows(o, r, rr);
            load_rows(o, dinv, dd);
            load_weights(o, wr, bx, by, bz);
            if (COARSE) {
    #pragma unroll
                for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                    for (int dy = 0; dy < 2; dy++) er[dz][dy] = ldr4(react, o.row[dz][dy]);
            }
            // coarse corrections: one aggregate pair per row group (the brick's rows share parents)
            const uint32_t cx = (uint32_t)(o.x0 >> 1), cy = (uint32_t)(o.y0 >> 1), cz = (uint32_t)(o.z0 >> 1);
            R2<VT> zcore = ldr2(zc, cx + csy * cy + csz * cz);
            R2<VT> zym = ldr2(zc, cx + csy * (uint32_t)(o.ymy >> 1) + csz * cz);
            R2<VT> zyp = ldr2(zc, cx + csy * (uint32_t)(o.ypy >> 1) + csz * cz);
            R2<VT> zzm = ldr2(zc, cx + csy * cy + csz * (uint32_t)(o.zmz >> 1));
            R2<VT> zzp = ldr2(zc, cx + csy * cy + csz * (uint32_t)(o.zpz >> 1));
            R1<VT> zxm = ldr1(zc, (uint32_t)(o.xmx >> 1) + csy * cy + csz * cz);
            R1<VT> zxp = ldr1(zc, (uint32_t)(o.xpx >> 1) + csy * cy + csz * cz);
            fence_rows(rr);
            fence_rows(dd);
            fence_weights(wr);
            if (COARSE) {
    #pragma unroll
                for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                    for (int dy = 0; dy < 2; dy++) fence(er[dz][dy]);
            }
            fence(zcore);
            fence(zym);
            fence(zyp);
            fence(zzm);
            fence(zzp);
            fence(zxm);
            fence(zxp);
            float cc[5][2];
            un(zcore, cc[0][0], cc[0][1]);
            un(zym, cc[1][0], cc[1][1]);
            un(zyp, cc[2][0], cc[2][1]);
            un(zzm, cc[3][0], cc[3][1]);
            un(zzp, cc[4][0], cc[4][1]);
            const float cxm = un(zxm), cxp = un(zxp);
            typename WtsOf<CT>::type w;
            weights(o, wr, w);
            Field f;
            field2(
                rr, dd, f,
                [&](const F4& a, const F4& d, int id) {
                    const int grp = id < 4 ? 0 : (id < 6 ? 1 : (id < 8 ? 2 : (id < 10 ? 3 : 4)));
                    const float c0 = cc[grp][0], c1 = cc[grp][1];
                    F4 o4;
    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        o4.v[k] = d.v[k] != 0.0f ? omega * (a.v[k] - m) * d.v[k] + ps * (k < 2 ? c0 : c1) : 0.0f;
                    return o4;
                },
                [&](float a, float d, int id) {
                    const float v = omega * (a - m) * d + ps * (id < 4 ? cxm : cxp);
                    return d != 0.0f ? v : 0.0f;
                });
    #pragma unroll
            for (int dz = 0; dz < 2; dz++)
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const uint32_t row = o.row[dz][dy];
                    const F4 rv = un(rr.c[dz][dy]), di = un(dd.c[dz][dy]);
                    F4 re{{0.f, 0.f, 0.f, 0.f}};
                    if (COARSE) re = un(er[dz][dy]);
                    F4 tv;
    #pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float rk = rv.v[k] - m;
                        float a = act_at(w, f, dz, dy, k);
                        if (COARSE) a = re.v[k] * f.c[dz][dy][k] + a;
                        const float out = f.c[dz][dy][k] + omega * (rk - a) * di.v[k];
                        const bool on = di.v[k] != 0.0f;
                        tv.v[k] = on ? out : 0.0f;
                        if (!COARSE) acc[0] += on ? (double)rk * (double)out : 0.0;
                    }
                    if (COARSE && A_.base) {  // each thread reads and writes only its own cells
                        const F4 bv = un(ldr4(A_.base, row));
#pragma unroll
                        for (int k = 0; k < 4; k++) tv.v[k] = bv.v[k] + A_.theta * tv.v[k];
                    }
                    st4(t, row, tv);
                }
        }
        if constexpr (!COARSE) reduce_finish<1>(vg, acc, part, sc, 4, mode);
    }
};

// r = r - A z on a coarse level (Galerkin graph with reaction term), in place: the residual one
// visit of the level leaves, the next visit's right-hand side.  Each thread reads its own r only.
template <class CT, class VT, int MB_>
struct K_k_brick_coarse_residual {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 g;
        const CT* bx;
        const CT* by;
        const CT* bz;
        const CT* react;
        const VT* z;
        VT* r;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        if (CONVERGED(A_.sc)) return;
        const Grid3 g = A_.g;
        uint32_t lo, hi;
        block_range(vg, g.n >> 4, lo, hi);
        for (uint32_t b = lo + threadIdx.x; b < hi; b += MHDP_BLOCK) {
            const Geo o = geo(g, b);
            Rows<VT> zr;
            WRaw<CT> wr;
            R4<CT> er[2][2];
            R4<VT> rc[2][2];
            load_rows(o, A_.z, zr);
            load_weights(o, wr, A_.bx, A_.by, A_.bz);
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    er[dz][dy] = ldr4(A_.react, o.row[dz][dy]);
                    rc[dz][dy] = ldr4(A_.r, o.row[dz][dy]);
                }
            fence_rows(zr);
            fence_weights(wr);
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    fence(er[dz][dy]);
                    fence(rc[dz][dy]);
                }
            typename WtsOf<CT>::type w;
            weights(o, wr, w);
            Field f;
            field2(
                zr, zr, f, [&](const F4& a, const F4&, int) { return a; }, [&](float a, float, int) { return a; });
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const F4 re = un(er[dz][dy]), rv = un(rc[dz][dy]);
                    F4 out;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float a = re.v[k] * f.c[dz][dy][k] + act_at(w, f, dz, dy, k);
                        out.v[k] = rv.v[k] - a;
                    }
                    st4(A_.r, o.row[dz][dy], out);
                }
        }
    }
};

// ---------------------------------------------------------------- launchers
// Minimum resident blocks per CU the register allocation targets (launch_tuning: 2, 3, 4).
static int brick_occ() { return launch_tuning().brick_occ; }
#define BRICK_DISPATCH(KERNEL, TARGS, ...)                                                          \
    do {                                                                                          \
        switch (brick_occ()) {                                                                    \
        case 4: go<K_##KERNEL<TARGS, 4>>(s, G.x, {__VA_ARGS__}); break;             \
        case 3: go<K_##KERNEL<TARGS, 3>>(s, G.x, {__VA_ARGS__}); break;             \
        default: go<K_##KERNEL<TARGS, 2>>(s, G.x, {__VA_ARGS__}); break;            \
        }                                                                                         \
    } while (0)

void launch_brick_apply(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                        const float* bx, const float* by, const float* bz, const void* dinv, float* q, double* part,
                        Scal* sc, int meanfree) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16)
        BRICK_DISPATCH(k_brick_apply, __hip_bfloat16, g, (const __hip_bfloat16*)z, p0, p1, x, bx, by, bz,
                       (const __hip_bfloat16*)dinv, q, part, sc, meanfree);
    else
        BRICK_DISPATCH(k_brick_apply, float, g, (const float*)z, p0, p1, x, bx, by, bz, (const float*)dinv, q, part, sc,
                       meanfree);
}
// launch_brick_apply with the mobilities formed from the liquid fraction (null: single phase,
// every open face 1) and the open-face bytes
void launch_brick_apply_m(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                          const float* alpha, const uint8_t* open8, float rho_l, float rho_air, float* q,
                          double* part, Scal* sc, int meanfree) {
    if (launch_apply_march(s, g, store, z, p0, p1, x, alpha, open8, rho_l, rho_air, q, part, sc, meanfree)) return;
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    const MobP mp{rho_l, rho_air};
#define COMMA ,
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        if (alpha)
            BRICK_DISPATCH(k_brick_apply_m, T COMMA true, g, (const T*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree);
        else
            BRICK_DISPATCH(k_brick_apply_m, T COMMA false, g, (const T*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree);
    } else {
        using T = float;
        if (alpha)
            BRICK_DISPATCH(k_brick_apply_m, T COMMA true, g, (const T*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree);
        else
            BRICK_DISPATCH(k_brick_apply_m, T COMMA false, g, (const T*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree);
    }
}
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
void launch_brick_smooth(hipStream_t s, int store, const Grid3& g, const Grid3& cg, const float* r, const float* r2,
                         const void* dinv, const void* bx, const void* by, const void* bz, const void* zc, void* t,
                         double* part, Scal* sc, int mode, bool use_next) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    const int un = use_next ? 1 : 0;
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA float COMMA false, g, cg, r, r2, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)nullptr, (const T*)zc, (T*)t, part, sc, mode, un);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA float COMMA false, g, cg, r, r2, (const T*)dinv, (const T*)bx,
                       (const T*)by, (const T*)bz, (const T*)nullptr, (const T*)zc, (T*)t, part, sc, mode, un);
    }
}
void launch_brick_restrict_update(hipStream_t s, int store, const Grid3& cg, const Grid3& g, float* r, float* r2,
                                  const float* q, const void* dinv, const void* bx, const void* by, const void* bz,
                                  void* rc, double* part, Scal* sc) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_restrict_update, T COMMA T, cg, g, r, r2, q, (const T*)dinv, (const T*)bx, (const T*)by,
                       (const T*)bz, (T*)rc, part, sc);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_restrict_update, T COMMA T, cg, g, r, r2, q, (const T*)dinv, (const T*)bx, (const T*)by,
                       (const T*)bz, (T*)rc, part, sc);
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
                                const void* zc, void* t, Scal* sc, const void* base, float theta) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA T COMMA true, g, cg, (const T*)r, (const T*)r, (const T*)dinv,
                       (const T*)bx, (const T*)by, (const T*)bz, (const T*)react, (const T*)zc, (T*)t, (double*)nullptr,
                       sc, 0, 0, (const T*)base, theta);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_smooth, T COMMA T COMMA T COMMA true, g, cg, (const T*)r, (const T*)r, (const T*)dinv,
                       (const T*)bx, (const T*)by, (const T*)bz, (const T*)react, (const T*)zc, (T*)t, (double*)nullptr,
                       sc, 0, 0, (const T*)base, theta);
    }
}
void launch_brick_coarse_residual(hipStream_t s, int store, const Grid3& g, const void* bx, const void* by,
                                  const void* bz, const void* react, const void* z, void* r, Scal* sc) {
    const dim3 B(MHDP_BLOCK), G(grid_blocks(g.n >> 4, brick_grid_cap()));
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        BRICK_DISPATCH(k_brick_coarse_residual, T COMMA T, g, (const T*)bx, (const T*)by, (const T*)bz,
                       (const T*)react, (const T*)z, (T*)r, sc);
    } else {
        using T = float;
        BRICK_DISPATCH(k_brick_coarse_residual, T COMMA T, g, (const T*)bx, (const T*)by, (const T*)bz,
                       (const T*)react, (const T*)z, (T*)r, sc);
    }
}
