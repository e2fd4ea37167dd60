// Z-marching forms of the fused fine-grid passes.  Each thread owns a column of 4 x 2 cells
// (four along x, two rows) and walks it up one z-chunk, one plane per step:
//   * the planes below and above the current one are held in registers (a window of three
//     planes), so every cell value is fetched from memory once per pass instead of once per
//     brick that touches it (the 4x2x2 bricks fetch the planes above and below each brick
//     again, ~1.3-1.6x the algorithmic bytes);
//   * the loads of plane k+2 are issued before plane k is computed (one plane of prefetch), so
//     a wave always has a plane of loads in flight while it computes.
// In-plane neighbours come from the same loads as the bricks (the rows y0-1, y0+2 and the
// cells x0-1, x0+4).  The arithmetic and summation order per cell are those of brick.hip.
// Needs nx % 4 == 0 and even ny, nz; a chunk is an even number of planes.
#include "mhdp_kernels.h"
#include "launch.h"
#include "raw.h"

#define CONVERGED(sc) ((sc)->conv > 0.5)

namespace {

// Column geometry (plane-independent parts).
struct Col {
    int x0, y0;
    uint32_t row0;           // x0 + nx*y0
    uint32_t ymr, ypr;       // starts of the rows y0-1 and y0+2 (wrapped; clamped when closed)
    int xm_off, xp_off;      // cells x0-1 and x0+4 relative to the row start
    int xmx, xpx, ymy, ypy;  // their coordinates
    bool xp_ok, yp_ok;       // the high-side faces of the column are open to a neighbour
};
__device__ __forceinline__ Col column(const Grid3& g, uint32_t cx, uint32_t cy) {
    Col o;
    o.x0 = (int)(cx << 2);
    o.y0 = (int)(cy << 1);
    const uint32_t sy = (uint32_t)g.nx;
    o.row0 = (uint32_t)o.x0 + sy * (uint32_t)o.y0;
    if (o.y0 > 0) o.ymy = o.y0 - 1;
    else o.ymy = (g.per & 2u) ? g.ny - 1 : o.y0;
    if (o.y0 + 2 < g.ny) { o.ypy = o.y0 + 2; o.yp_ok = true; }
    else if (g.per & 2u) { o.ypy = 0; o.yp_ok = true; }
    else { o.ypy = o.y0 + 1; o.yp_ok = false; }
    o.ymr = (uint32_t)o.x0 + sy * (uint32_t)o.ymy;
    o.ypr = (uint32_t)o.x0 + sy * (uint32_t)o.ypy;
    if (o.x0 > 0) o.xmx = o.x0 - 1;
    else o.xmx = (g.per & 1u) ? g.nx - 1 : o.x0;
    if (o.x0 + 4 < g.nx) { o.xpx = o.x0 + 4; o.xp_ok = true; }
    else if (g.per & 1u) { o.xpx = 0; o.xp_ok = true; }
    else { o.xpx = o.x0 + 3; o.xp_ok = false; }
    o.xm_off = o.xmx - o.x0;
    o.xp_off = o.xpx - o.x0;
    return o;
}
// Plane p in [-1, nz]: wrapped when periodic, clamped otherwise (its faces are closed).
__device__ __forceinline__ int plane_of(const Grid3& g, int p) {
    if (p < 0) return (g.per & 4u) ? p + g.nz : 0;
    if (p >= g.nz) return (g.per & 4u) ? p - g.nz : g.nz - 1;
    return p;
}

// A field on one plane of the column and its in-plane halo.
struct PF {
    float c[2][4];          // [dy][k]
    float ym[4], yp[4];     // rows y0-1, y0+2
    float xm[2], xp[2];     // cells x0-1, x0+4 of rows y0, y0+1
};
// One array on one plane of the column: own rows, halo rows, halo cells.
template <class T>
struct PR {
    R4<T> c[2], ym, yp;
    R1<T> xm[2], xp[2];
};
template <class T>
__device__ __forceinline__ void load_pr(const Col& o, uint32_t base, uint32_t sy, const T* __restrict__ p, PR<T>& r) {
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
        const uint32_t row = base + o.row0 + (uint32_t)dy * sy;
        r.c[dy] = ldr4(p, row);
        r.xm[dy] = ldr1(p, row + (uint32_t)o.xm_off);
        r.xp[dy] = ldr1(p, row + (uint32_t)o.xp_off);
    }
    r.ym = ldr4(p, base + o.ymr);
    r.yp = ldr4(p, base + o.ypr);
}
template <class T>
__device__ __forceinline__ void fence_pr(PR<T>& r) {
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
        fence(r.c[dy]);
        fence(r.xm[dy]);
        fence(r.xp[dy]);
    }
    fence(r.ym);
    fence(r.yp);
}
// Mobilities of one plane: lower x/y/z faces of the own cells, the x+ face of each row's last
// cell, and the lower y face of row y0+2 (= the upper y face of row y0+1).
template <class T>
struct PW {
    R4<T> bx[2], by[2], bz[2], byp;
    R1<T> bxp[2];
};
template <class T>
__device__ __forceinline__ void load_pw(const Col& o, uint32_t base, uint32_t sy, const T* __restrict__ bx,
                                        const T* __restrict__ by, const T* __restrict__ bz, PW<T>& w) {
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
        const uint32_t row = base + o.row0 + (uint32_t)dy * sy;
        w.bx[dy] = ldr4(bx, row);
        w.by[dy] = ldr4(by, row);
        w.bz[dy] = ldr4(bz, row);
        w.bxp[dy] = ldr1(bx, row + (uint32_t)o.xp_off);
    }
    w.byp = ldr4(by, base + o.ypr);
}
template <class T>
__device__ __forceinline__ void fence_pw(PW<T>& w) {
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
        fence(w.bx[dy]);
        fence(w.by[dy]);
        fence(w.bz[dy]);
        fence(w.bxp[dy]);
    }
    fence(w.byp);
}
// Six face weights of own cell (dy, k) in the order x-, x+, y-, y+, z-, z+; `up` holds the
// lower z faces of the plane above (closed at a non-periodic top: zp_ok false).
template <class T>
__device__ __forceinline__ void face6(const Col& o, const PW<T>& w, const PW<T>& up, bool zp_ok, int dy, int k,
                                      float (&f)[6]) {
    const F4 bxr = un(w.bx[dy]), byr = un(w.by[dy]), bzr = un(w.bz[dy]), bzu = un(up.bz[dy]);
    f[0] = bxr.v[k];
    f[1] = k < 3 ? bxr.v[k + 1] : (o.xp_ok ? un(w.bxp[dy]) : 0.0f);
    f[2] = byr.v[k];
    f[3] = dy == 0 ? un(w.by[1]).v[k] : (o.yp_ok ? un(w.byp).v[k] : 0.0f);
    f[4] = bzr.v[k];
    f[5] = zp_ok ? bzu.v[k] : 0.0f;
}
// A v at own cell (dy, k) of plane k: in-plane neighbours from `c`, z neighbours from the
// planes below and above.
__device__ __forceinline__ float act6(const PF& c, const float (&below)[2][4], const float (&above)[2][4], int dy,
                                      int k, const float (&w)[6]) {
    const float v = c.c[dy][k];
    const float xm = k > 0 ? c.c[dy][k - 1] : c.xm[dy];
    const float xp = k < 3 ? c.c[dy][k + 1] : c.xp[dy];
    const float ym = dy == 1 ? c.c[0][k] : c.ym[k];
    const float yp = dy == 0 ? c.c[1][k] : c.yp[k];
    float a = 0.0f;
    a += w[0] * (v - xm);
    a += w[1] * (v - xp);
    a += w[2] * (v - ym);
    a += w[3] * (v - yp);
    a += w[4] * (v - below[dy][k]);
    a += w[5] * (v - above[dy][k]);
    return a;
}
// Work items: columns (x fastest) of one z-chunk, chunks slowest.
struct Item {
    uint32_t cx, cy;
    int za, zb;
};
__device__ __forceinline__ Item item_of(const Grid3& g, uint32_t it, uint32_t zc) {
    const uint32_t CX = (uint32_t)g.nx >> 2, C = CX * ((uint32_t)g.ny >> 1);
    const uint32_t ch = it / C, j = it - ch * C;
    Item o;
    o.cy = j / CX;
    o.cx = j - o.cy * CX;
    o.za = (int)(ch * zc);
    o.zb = min(o.za + (int)zc, g.nz);
    return o;
}

}  // namespace

// Pass 4 (post-smoother): z2 = omega (r - m)/d + ps * zc[parent] on active cells (prolongation
// and pre-smoother, never stored), t = z2 + omega (r - m - A z2)/d, and (r - m).t.
template <class CT, class VT, int MB_>
struct K_k_march_smooth {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 g;
        Grid3 cg;
        const float* r0;
        const float* r1;
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
        const VT* zc;
        VT* t;
        double* part;
        Scal* sc;
        int mode;
        int use_next;
        uint32_t zlen;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        Grid3 cg = A_.cg;
        const float* r0 = A_.r0;
        const float* r1 = A_.r1;
        const CT* __restrict__ dinv = A_.dinv;
        const CT* __restrict__ bx = A_.bx;
        const CT* __restrict__ by = A_.by;
        const CT* __restrict__ bz = A_.bz;
        const VT* __restrict__ zc = A_.zc;
        VT* __restrict__ t = A_.t;
        double* __restrict__ part = A_.part;
        Scal* sc = A_.sc;
        int mode = A_.mode;
        int use_next = A_.use_next;
        uint32_t zlen = A_.zlen;
        (void)vg;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;
        const float m = (float)(use_next ? sc->mean_next : sc->mean);
        const float ps = (float)sc->pscale;
        const float* __restrict__ r = sc->rsel ? r1 : r0;
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const uint32_t csy = (uint32_t)cg.nx, csz = (uint32_t)cg.nx * (uint32_t)cg.ny;
        const uint32_t items = ((uint32_t)g.nx >> 2) * ((uint32_t)g.ny >> 1) * (((uint32_t)g.nz + zlen - 1) / zlen);
        double acc[1] = {0.0};
        for (uint32_t it = vg.b * MHDP_BLOCK + threadIdx.x; it < items; it += vg.n * MHDP_BLOCK) {
            const Item w0 = item_of(g, it, zlen);
            const Col o = column(g, w0.cx, w0.cy);
            const uint32_t cx = w0.cx << 1;  // parents of cells x0, x0+1 and x0+2, x0+3: cx, cx+1
            const uint32_t cy = (uint32_t)o.y0 >> 1, cym = (uint32_t)o.ymy >> 1, cyp = (uint32_t)o.ypy >> 1;
            const uint32_t cxm = (uint32_t)o.xmx >> 1, cxp = (uint32_t)o.xpx >> 1;
            struct Raw {
                PR<float> r;
                PR<CT> d;
                R2<VT> zo, zym, zyp;
                R1<VT> zxm, zxp;
                PW<CT> w;
            };
            auto load = [&](int p, Raw& q) {
                const int pp = plane_of(g, p);
                const uint32_t base = sz * (uint32_t)pp, cz = csz * ((uint32_t)pp >> 1);
                load_pr(o, base, sy, r, q.r);
                load_pr(o, base, sy, dinv, q.d);
                load_pw(o, base, sy, bx, by, bz, q.w);
                q.zo = ldr2(zc, cz + csy * cy + cx);
                q.zym = ldr2(zc, cz + csy * cym + cx);
                q.zyp = ldr2(zc, cz + csy * cyp + cx);
                q.zxm = ldr1(zc, cz + csy * cy + cxm);
                q.zxp = ldr1(zc, cz + csy * cy + cxp);
            };
            auto settle = [&](Raw& q) {
                fence_pr(q.r);
                fence_pr(q.d);
                fence_pw(q.w);
                fence(q.zo);
                fence(q.zym);
                fence(q.zyp);
                fence(q.zxm);
                fence(q.zxp);
            };
            // z2 on the plane's own cells and in-plane halo
            auto field = [&](const Raw& q, PF& f) {
                float c0, c1, y0a, y0b, y1a, y1b;
                un(q.zo, c0, c1);
                un(q.zym, y0a, y0b);
                un(q.zyp, y1a, y1b);
                const float cxmv = un(q.zxm), cxpv = un(q.zxp);
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const F4 a = un(q.r.c[dy]), d = un(q.d.c[dy]);
    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        f.c[dy][k] = d.v[k] != 0.0f ? omega * (a.v[k] - m) * d.v[k] + ps * (k < 2 ? c0 : c1) : 0.0f;
                    const float dm = un(q.d.xm[dy]), dp = un(q.d.xp[dy]);
                    const float vm = omega * (un(q.r.xm[dy]) - m) * dm + ps * cxmv;
                    const float vp = omega * (un(q.r.xp[dy]) - m) * dp + ps * cxpv;
                    f.xm[dy] = dm != 0.0f ? vm : 0.0f;
                    f.xp[dy] = dp != 0.0f ? vp : 0.0f;
                }
                const F4 am = un(q.r.ym), dm = un(q.d.ym), ap = un(q.r.yp), dp = un(q.d.yp);
    #pragma unroll
                for (int k = 0; k < 4; k++) {
                    f.ym[k] = dm.v[k] != 0.0f ? omega * (am.v[k] - m) * dm.v[k] + ps * (k < 2 ? y0a : y0b) : 0.0f;
                    f.yp[k] = dp.v[k] != 0.0f ? omega * (ap.v[k] - m) * dp.v[k] + ps * (k < 2 ? y1a : y1b) : 0.0f;
                }
            };
            Raw cur, nxt;
            PF fc, fn;
            float below[2][4];
            {
                Raw prev;
                load(w0.za - 1, prev);
                load(w0.za, cur);
                load(w0.za + 1, nxt);
                settle(prev);
                settle(cur);
                PF fp;
                field(prev, fp);
    #pragma unroll
                for (int dy = 0; dy < 2; dy++)
    #pragma unroll
                    for (int k = 0; k < 4; k++) below[dy][k] = fp.c[dy][k];
                field(cur, fc);
            }
            for (int k = w0.za; k < w0.zb; k++) {
                Raw pre;
                load(k + 2, pre);
                settle(nxt);
                field(nxt, fn);
                const bool zp_ok = k + 1 < g.nz || (g.per & 4u);
                const uint32_t base = sz * (uint32_t)k;
    #pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const F4 rv = un(cur.r.c[dy]), di = un(cur.d.c[dy]);
                    F4 tv;
    #pragma unroll
                    for (int kk = 0; kk < 4; kk++) {
                        float fw[6];
                        face6(o, cur.w, nxt.w, zp_ok, dy, kk, fw);
                        const float rk = rv.v[kk] - m;
                        const float a = act6(fc, below, fn.c, dy, kk, fw);
                        const float out = fc.c[dy][kk] + omega * (rk - a) * di.v[kk];
                        const bool on = di.v[kk] != 0.0f;
                        tv.v[kk] = on ? out : 0.0f;
                        acc[0] += on ? (double)rk * (double)out : 0.0;
                    }
                    st4(t, base + o.row0 + (uint32_t)dy * sy, tv);
                }
    #pragma unroll
                for (int dy = 0; dy < 2; dy++)
    #pragma unroll
                    for (int kk = 0; kk < 4; kk++) below[dy][kk] = fc.c[dy][kk];
                fc = fn;
                cur = nxt;
                nxt = pre;
            }
        }
        reduce_finish<1>(vg, acc, part, sc, 4, mode);
    }
};

// Passes 2+3 (restriction with the residual update) on z-marching columns: r' = (r - m) - alpha q
// on the column's planes and their halos, r' stored for the own cells in the twin buffer, the
// pre-smoothed z1 = omega (r' - m') / d, the residual (r' - m') - A z1 summed over each 2x2x2
// aggregate (two per column and plane pair), and sum r', sum r'^2 -> FIN_RESTRICT_UPD.
// Arithmetic per cell as K_k_brick_restrict_update.
template <class CT, class VT, int MB_>
struct K_march_restrict_update {
    static constexpr int MINB = MB_;
    struct Args {
        Grid3 cg, g;
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
        uint32_t zlen;
    };
    struct Raw {
        PR<float> r, q;
        PR<CT> d;
        PW<CT> w;
    };
    // r' and z1 of one plane (own cells and in-plane halo); rown: r' of the own cells
    struct Fields {
        PF z1;
        float rown[2][4];
    };
    static __device__ void run(const Args& A, const VGrid& vg) {
        const Grid3 g = A.g;
        Scal* sc = A.sc;
        if (CONVERGED(sc)) return;
        const float omega = (float)sc->omega;
        const float alpha = (float)sc->alpha;
        const float m = (float)sc->mean;
        const float mn = (float)sc->mean_next;
        const float* __restrict__ rin = sc->rsel ? A.r1 : A.r0;
        float* __restrict__ rout = sc->rsel ? A.r0 : A.r1;
        const float* __restrict__ q = A.q;
        const CT* __restrict__ dinv = A.dinv;
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const uint32_t csy = (uint32_t)A.cg.nx, csz = (uint32_t)A.cg.nx * (uint32_t)A.cg.ny;
        const uint32_t zlen = A.zlen;
        const uint32_t items = ((uint32_t)g.nx >> 2) * ((uint32_t)g.ny >> 1) * (((uint32_t)g.nz + zlen - 1) / zlen);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t it = vg.b * MHDP_BLOCK + threadIdx.x; it < items; it += vg.n * MHDP_BLOCK) {
            const Item w0 = item_of(g, it, zlen);
            const Col o = column(g, w0.cx, w0.cy);
            auto load = [&](int p, Raw& x) {
                const uint32_t base = sz * (uint32_t)plane_of(g, p);
                load_pr(o, base, sy, rin, x.r);
                load_pr(o, base, sy, q, x.q);
                load_pr(o, base, sy, dinv, x.d);
                load_pw(o, base, sy, A.bx, A.by, A.bz, x.w);
            };
            auto settle = [&](Raw& x) {
                fence_pr(x.r);
                fence_pr(x.q);
                fence_pr(x.d);
                fence_pw(x.w);
            };
            auto upd = [&](float r, float qq, float d) { return d != 0.0f ? (r - m) - alpha * qq : 0.0f; };
            auto field = [&](const Raw& x, Fields& f) {
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const F4 a = un(x.r.c[dy]), b = un(x.q.c[dy]), d = un(x.d.c[dy]);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float rn = upd(a.v[k], b.v[k], d.v[k]);
                        f.rown[dy][k] = rn;
                        f.z1.c[dy][k] = omega * (rn - mn) * d.v[k];
                    }
                    const float dm = un(x.d.xm[dy]), dp = un(x.d.xp[dy]);
                    f.z1.xm[dy] = omega * (upd(un(x.r.xm[dy]), un(x.q.xm[dy]), dm) - mn) * dm;
                    f.z1.xp[dy] = omega * (upd(un(x.r.xp[dy]), un(x.q.xp[dy]), dp) - mn) * dp;
                }
                const F4 am = un(x.r.ym), bm = un(x.q.ym), dm = un(x.d.ym);
                const F4 ap = un(x.r.yp), bp = un(x.q.yp), dp = un(x.d.yp);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    f.z1.ym[k] = omega * (upd(am.v[k], bm.v[k], dm.v[k]) - mn) * dm.v[k];
                    f.z1.yp[k] = omega * (upd(ap.v[k], bp.v[k], dp.v[k]) - mn) * dp.v[k];
                }
            };
            Raw cur, nxt;
            Fields fc, fn;
            float below[2][4];
            {
                Raw prev;
                load(w0.za - 1, prev);
                load(w0.za, cur);
                load(w0.za + 1, nxt);
                settle(prev);
                settle(cur);
                Fields fp;
                field(prev, fp);
#pragma unroll
                for (int dy = 0; dy < 2; dy++)
#pragma unroll
                    for (int k = 0; k < 4; k++) below[dy][k] = fp.z1.c[dy][k];
                field(cur, fc);
            }
            float s0 = 0.0f, s1 = 0.0f;
            for (int k = w0.za; k < w0.zb; k++) {
                Raw pre;
                load(k + 2, pre);
                settle(nxt);
                field(nxt, fn);
                const bool zp_ok = k + 1 < g.nz || (g.per & 4u);
                const uint32_t base = sz * (uint32_t)k;
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const F4 di = un(cur.d.c[dy]);
                    F4 ro;
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) {
                        const float rk = fc.rown[dy][kk];
                        ro.v[kk] = rk;
                        acc[1] += (double)rk * (double)rk;
                        acc[3] += (double)rk;
                        float fw[6];
                        face6(o, cur.w, nxt.w, zp_ok, dy, kk, fw);
                        const float a = act6(fc.z1, below, fn.z1.c, dy, kk, fw);
                        float v = (rk - mn) - a;
                        v = di.v[kk] != 0.0f ? v : 0.0f;
                        if (kk < 2) s0 += v;
                        else s1 += v;
                    }
                    st4(rout, base + o.row0 + (uint32_t)dy * sy, ro);
                }
                if (k & 1) {  // the second plane of an aggregate pair
                    const uint32_t ci = (uint32_t)(o.x0 >> 1) + csy * (uint32_t)(o.y0 >> 1) + csz * (uint32_t)(k >> 1);
                    st2(A.rc, ci, s0, s1);
                    s0 = 0.0f;
                    s1 = 0.0f;
                }
#pragma unroll
                for (int dy = 0; dy < 2; dy++)
#pragma unroll
                    for (int kk = 0; kk < 4; kk++) below[dy][kk] = fc.z1.c[dy][kk];
                fc = fn;
                cur = nxt;
                nxt = pre;
            }
        }
        reduce_finish<4>(vg, acc, A.part, sc, 3, FIN_RESTRICT_UPD);
    }
};

// ---------------------------------------------------------------- launchers

// Planes per work item and the block cap: launch_tuning() (defaults 32 and 1024).
static uint32_t march_z() { return launch_tuning().march_z; }
bool march_ok(const Grid3& g) {
    static const bool on = [] {
        const char* e = std::getenv("MHDP_NO_MARCH");
        return !(e && e[0] == '1');
    }();
    if (!on || (g.nx & 3) != 0 || (g.ny & 1) != 0 || (g.nz & 1) != 0 || g.nz < 2) return false;
    // a chunk is walked by one thread: small grids have too few columns to fill the device and
    // keep the 4x2x2 bricks (one thread per 16 cells).  MHDP_MARCH_MIN (work items, read per
    // call) moves the threshold, e.g. 0 to exercise this path on small grids.
    const char* e = std::getenv("MHDP_MARCH_MIN");
    const uint64_t min_items = e ? (uint64_t)std::strtoull(e, nullptr, 10) : 65536u;
    const uint32_t zl = march_z();
    const uint64_t items = (uint64_t)((uint32_t)g.nx >> 2) * ((uint32_t)g.ny >> 1) * (((uint32_t)g.nz + zl - 1) / zl);
    return items >= min_items;
}
static uint32_t march_blocks(const Grid3& g) {
    const uint32_t cap = launch_tuning().march_cap;
    const uint32_t zl = march_z();
    const uint32_t items = ((uint32_t)g.nx >> 2) * ((uint32_t)g.ny >> 1) * (((uint32_t)g.nz + zl - 1) / zl);
    return grid_blocks(items, cap);
}

void launch_march_smooth(hipStream_t s, const Grid3& g, const Grid3& cg, const float* r, const float* r2,
                         const void* dinv, const void* bx, const void* by, const void* bz, const void* zc, void* t,
                         double* part, Scal* sc, int mode, bool use_next) {
    using T = __hip_bfloat16;
    go<K_k_march_smooth<T, T, 2>>(s, march_blocks(g), {g, cg, r, r2, (const T*)dinv, (const T*)bx, (const T*)by,
                                                       (const T*)bz, (const T*)zc, (T*)t, part, sc, mode,
                                                       use_next ? 1 : 0, march_z()});
}

void launch_march_restrict_update(hipStream_t s, const Grid3& cg, const Grid3& g, float* r, float* r2,
                                  const float* q, const void* dinv, const void* bx, const void* by, const void* bz,
                                  void* rc, double* part, Scal* sc) {
    using T = __hip_bfloat16;
    go<K_march_restrict_update<T, T, 2>>(s, march_blocks(g), {cg, g, r, r2, q, (const T*)dinv, (const T*)bx,
                                                              (const T*)by, (const T*)bz, (T*)rc, part, sc,
                                                              march_z()});
}
