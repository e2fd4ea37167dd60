// Once-per-projection kernels: face mobilities + diagonal, divergence RHS, velocity
// correction and the Galerkin coarse graphs.  Built with -ffp-contract=off: their FP32
// results are bit-identical to the reference formulas evaluated in IEEE order.
#include "mhdp_kernels.h"
#include "launch.h"
#include "rows.h"

namespace {

__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

struct Pos {
    int p[3];
};

__device__ __forceinline__ bool canon_valid(const Grid3& g, Pos& q) {
    const int d[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (g.per & (1u << a)) q.p[a] = wrap1(q.p[a], d[a]);
        if (q.p[a] < 0 || q.p[a] >= d[a]) return false;
    }
    return true;
}
__device__ __forceinline__ uint32_t cidx(const Grid3& g, const Pos& q) {
    return (uint32_t)q.p[0] + (uint32_t)g.nx * ((uint32_t)q.p[1] + (uint32_t)g.ny * (uint32_t)q.p[2]);
}
// fluid(), solver.wgsl:29
__device__ __forceinline__ bool fluid_at(const Grid3& g, const uint32_t* __restrict__ mask, Pos q) {
    if (!canon_valid(g, q)) return false;
    return mask[cidx(g, q)] != 0u;
}
// occupancy()/density(), solver.wgsl:40-41
__device__ __forceinline__ float density_at(const Grid3& g, const PhaseP& ph, const float* __restrict__ alpha, Pos q) {
    float a = 0.0f;
    if (canon_valid(g, q)) a = alpha ? alpha[cidx(g, q)] : 0.0f;
    return ph.rho_air * (1.0f - a) + ph.rho_l * a;
}
// mobility(), solver.wgsl:43-51 (mushy factor off)
__device__ __forceinline__ float mobility_at(const Grid3& g, const PhaseP& ph, const float* __restrict__ alpha, Pos pos,
                                             int axis) {
    if (!ph.two_phase) return 1.0f;
    Pos left = pos;
    left.p[axis] -= 1;
    float fd = 0.5f * (density_at(g, ph, alpha, left) + density_at(g, ph, alpha, pos));
    return ph.rho_l / fd;
}

}  // namespace

// pressure_coefficients (solver.wgsl:93-103) in cell-indexed lower-face form, plus
// diagonal() (solver.wgsl:60-64) and the active-cell reciprocal.
struct K_k_coefficients {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        PhaseP ph;
        const uint32_t* mask;
        const float* alpha;
        float* bx;
        float* by;
        float* bz;
        float* diag;
        float* dinv;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        PhaseP ph = A_.ph;
        const uint32_t* __restrict__ mask = A_.mask;
        const float* __restrict__ alpha = A_.alpha;
        float* __restrict__ bx = A_.bx;
        float* __restrict__ by = A_.by;
        float* __restrict__ bz = A_.bz;
        float* __restrict__ diag = A_.diag;
        float* __restrict__ dinv = A_.dinv;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int x, y, z;
            decode(g, i, x, y, z);
            Pos pos{{x, y, z}};
            const uint32_t m = mask[i];
            const bool fc = m != 0u;
            float b[3];
    #pragma unroll
            for (int a = 0; a < 3; a++) {
                Pos left = pos;
                left.p[a] -= 1;
                b[a] = (fc && fluid_at(g, mask, left)) ? mobility_at(g, ph, alpha, pos, a) : 0.0f;
            }
            bx[i] = b[0];
            by[i] = b[1];
            bz[i] = b[2];
            float d = 0.0f;
            if (fc) {
    #pragma unroll
                for (int a = 0; a < 3; a++) {
    #pragma unroll
                    for (int s = -1; s <= 1; s += 2) {
                        Pos q = pos;
                        q.p[a] += s;
                        if (fluid_at(g, mask, q)) {
                            Pos fp = pos;
                            if (s > 0) fp.p[a] += 1;
                            d += mobility_at(g, ph, alpha, fp, a);
                        }
                    }
                }
                if (!ph.meanfree && (m & 3u) == 2u) d = 1.0f;
                else d = fmaxf(d, 1e-30f);
            }
            diag[i] = d;
            const bool active = fc && (ph.meanfree || (m & 3u) == 1u);
            dinv[i] = active ? 1.0f / d : 0.0f;
        }
    }
};

// Canonical face index in plane `a` of the upper face of cell (x,y,z) along a.
__device__ __forceinline__ uint32_t plane_index(const Grid3& g, int a, int fx, int fy, int fz) {
    const uint32_t Px = (uint32_t)g.nx + (a == 0), Py = (uint32_t)g.ny + (a == 1);
    return (uint32_t)fx + Px * ((uint32_t)fy + Py * (uint32_t)fz);
}

// rhs() without the pinned-anchor zeroing (solver.wgsl:104-107, 192): b = -h^2 div(u*)
// over open faces.  Openness is read from the mobilities (a face is open iff beta > 0).
template <class T>
struct K_k_rhs {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float h;
        const uint32_t* mask;
        const float* bx;
        const float* by;
        const float* bz;
        const float* ustar;
        uint64_t off1;
        uint64_t off2;
        T* b;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        float h = A_.h;
        const uint32_t* __restrict__ mask = A_.mask;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const float* __restrict__ ustar = A_.ustar;
        uint64_t off1 = A_.off1;
        uint64_t off2 = A_.off2;
        T* __restrict__ b = A_.b;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        const float* U[3] = {ustar, ustar + off1, ustar + off2};
        const float* B[3] = {bx, by, bz};
        const int d[3] = {g.nx, g.ny, g.nz};
        const uint32_t st[3] = {1u, (uint32_t)g.nx, (uint32_t)g.nx * (uint32_t)g.ny};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            T out = (T)0;
            if (mask[i] != 0u) {
                int x, y, z;
                decode(g, i, x, y, z);
                int c[3] = {x, y, z};
                T sum = (T)0;
    #pragma unroll
                for (int a = 0; a < 3; a++) {
                    // upper face: cell c + e_a (wrapped) lower face
                    float up = 0.0f;
                    int u[3] = {c[0], c[1], c[2]};
                    u[a] += 1;
                    bool exists = true;
                    uint32_t ui = i + st[a];
                    if (u[a] >= d[a]) {
                        if (g.per & (1u << a)) { u[a] = 0; ui = i - (uint32_t)(d[a] - 1) * st[a]; }
                        else exists = false;
                    }
                    if (exists && B[a][ui] > 0.0f) up = U[a][plane_index(g, a, u[0], u[1], u[2])];
                    float dn = 0.0f;
                    if (B[a][i] > 0.0f) dn = U[a][plane_index(g, a, c[0], c[1], c[2])];
                    sum += (T)up - (T)dn;
                }
                T div = sum / (T)h;
                out = -(T)h * (T)h * div;
            }
            b[i] = out;
        }
    }
};

// correct() (solver.wgsl:245-250) over every entry of the three face planes.
template <class PT>
struct K_k_correct {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float h;
        const float* bx;
        const float* by;
        const float* bz;
        const PT* phi;
        const float* ustar;
        float* uout;
        uint32_t f0;
        uint32_t f1;
        uint32_t f2;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        float h = A_.h;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const PT* __restrict__ phi = A_.phi;
        const float* __restrict__ ustar = A_.ustar;
        float* __restrict__ uout = A_.uout;
        uint32_t f0 = A_.f0;
        uint32_t f1 = A_.f1;
        uint32_t f2 = A_.f2;
        (void)vg;
        const uint32_t total = f0 + f1 + f2;
        uint32_t lo, hi;
        block_range(vg, total, lo, hi);
        const float* B[3] = {bx, by, bz};
        const int d[3] = {g.nx, g.ny, g.nz};
        const uint32_t st[3] = {1u, (uint32_t)g.nx, (uint32_t)g.nx * (uint32_t)g.ny};
        for (uint32_t f = lo + threadIdx.x; f < hi; f += MHDP_BLOCK) {
            int a;
            uint32_t k;
            if (f < f0) { a = 0; k = f; }
            else if (f < f0 + f1) { a = 1; k = f - f0; }
            else { a = 2; k = f - f0 - f1; }
            const uint32_t Px = (uint32_t)g.nx + (a == 0), Py = (uint32_t)g.ny + (a == 1);
            int q[3];
            q[0] = (int)(k % Px);
            q[1] = (int)((k / Px) % Py);
            q[2] = (int)(k / (Px * Py));
            float out = 0.0f;
            bool ok = true;
            if (q[a] >= d[a]) {
                if (g.per & (1u << a)) q[a] = 0;
                else ok = false;
            }
            if (ok) {
                const uint32_t ci = (uint32_t)q[0] + (uint32_t)g.nx * ((uint32_t)q[1] + (uint32_t)g.ny * (uint32_t)q[2]);
                const float beta = B[a][ci];
                if (beta > 0.0f) {
                    uint32_t li = q[a] > 0 ? ci - st[a] : ci + (uint32_t)(d[a] - 1) * st[a];
                    const float* U = ustar + (a == 0 ? 0 : (a == 1 ? f0 : f0 + f1));
                    const float uf = U[plane_index(g, a, q[0], q[1], q[2])];
                    if (sizeof(PT) == 4) {
                        out = uf - (float)beta * ((float)phi[ci] - (float)phi[li]) / h;
                    } else {
                        double o = (double)uf - (double)beta * ((double)phi[ci] - (double)phi[li]) / (double)h;
                        out = (float)o;
                    }
                }
            }
            uout[f] = out;
        }
    }
};

// build() (multigrid.wgsl:50-65): exact piecewise-constant Galerkin graph of one coarse
// level.  FINE: children are the fine cells (cached mobilities, active mask, anchors
// become reactions in the pinned gauge); otherwise children are the previous coarse level.
template <bool FINE>
struct K_k_mg_build {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 cg;
        Grid3 fg;
        int meanfree;
        const uint32_t* mask;
        const float* fdinv;
        const float* fbx;
        const float* fby;
        const float* fbz;
        const float* freact;
        float* cbx;
        float* cby;
        float* cbz;
        float* creact;
        float* cdiag;
        float* cdinv;
        __hip_bfloat16* c16[5];  // optional BF16 copies of bx, by, bz, react, dinv (null: none)
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 cg = A_.cg;
        Grid3 fg = A_.fg;
        int meanfree = A_.meanfree;
        const uint32_t* __restrict__ mask = A_.mask;
        const float* __restrict__ fdinv = A_.fdinv;
        const float* __restrict__ fbx = A_.fbx;
        const float* __restrict__ fby = A_.fby;
        const float* __restrict__ fbz = A_.fbz;
        const float* __restrict__ freact = A_.freact;
        float* __restrict__ cbx = A_.cbx;
        float* __restrict__ cby = A_.cby;
        float* __restrict__ cbz = A_.cbz;
        float* __restrict__ creact = A_.creact;
        float* __restrict__ cdiag = A_.cdiag;
        float* __restrict__ cdinv = A_.cdinv;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, cg.n, lo, hi);
        const float* FB[3] = {fbx, fby, fbz};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int qx, qy, qz;
            decode(cg, i, qx, qy, qz);
            const int q[3] = {qx, qy, qz};
            // every value the children and their six neighbours may need, loaded first from
            // always-valid indices (a child's own cell where a neighbour does not exist); the sums
            // below then run in the reference's order (children z, y, x; axes; lower then upper)
            bool cok[8], nok[8][6], inside[8][6];
            float dv[8], bt[8][6];
            uint32_t mk[8][6];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                Pos child{{2 * qx + (k & 1), 2 * qy + ((k >> 1) & 1), 2 * qz + (k >> 2)}};
                cok[k] = child.p[0] < fg.nx && child.p[1] < fg.ny && child.p[2] < fg.nz;
                const uint32_t ci = cok[k] ? cidx(fg, child) : 0u;
                dv[k] = FINE ? fdinv[ci] : freact[ci];
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        Pos n = child;
                        n.p[a] += h ? 1 : -1;
                        const bool ok = cok[k] && canon_valid(fg, n);
                        const uint32_t ni = ok ? cidx(fg, n) : ci;
                        nok[k][2 * a + h] = ok;
                        inside[k][2 * a + h] = n.p[0] / 2 == q[0] && n.p[1] / 2 == q[1] && n.p[2] / 2 == q[2];
                        bt[k][2 * a + h] = FB[a][h ? ni : ci];
                        mk[k][2 * a + h] = (FINE && !meanfree) ? mask[ni] : 0u;
                    }
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                asm volatile("" : "+v"(dv[k]));
#pragma unroll
                for (int f = 0; f < 6; f++) asm volatile("" : "+v"(bt[k][f]), "+v"(mk[k][f]));
            }
            float c4[4] = {0.f, 0.f, 0.f, 0.f};
            float diag = 0.f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (!cok[k]) continue;
                if (FINE) {
                    if (dv[k] == 0.0f) continue;
                } else {
                    c4[3] += dv[k];
                    diag += dv[k];
                }
#pragma unroll
                for (int f = 0; f < 6; f++) {
                    if (!nok[k][f]) continue;
                    const float beta = bt[k][f];
                    if (beta == 0.0f) continue;
                    if (FINE && !meanfree && (mk[k][f] & 3u) == 2u) {
                        c4[3] += beta;
                        diag += beta;
                        continue;
                    }
                    if (inside[k][f]) continue;
                    diag += beta;
                    if (!(f & 1)) c4[f >> 1] += beta;
                }
            }
            const float dinv = diag > 0.0f ? 1.0f / diag : 0.0f;
            cbx[i] = c4[0];
            cby[i] = c4[1];
            cbz[i] = c4[2];
            creact[i] = c4[3];
            cdiag[i] = diag;
            cdinv[i] = dinv;
            if (A_.c16[0]) {
                A_.c16[0][i] = __float2bfloat16(c4[0]);
                A_.c16[1][i] = __float2bfloat16(c4[1]);
                A_.c16[2][i] = __float2bfloat16(c4[2]);
                A_.c16[3][i] = __float2bfloat16(c4[3]);
                A_.c16[4][i] = __float2bfloat16(dinv);
            }
        }
    }
};

// K_k_mg_build<true> on fine grids whose axes are all even and at least 4 long.  Every aggregate
// then has its eight children, and no neighbour reached across a periodic wrap lies in the same
// aggregate, so a child's face is inside exactly when its neighbour is the sibling.  The
// children's values come as x pairs (8-byte loads) from their four rows, the outside neighbours
// from the rows around; indices are row bases plus offsets (a wrap only at the grid's ends).
// The sums run over the same values in the same order as K_k_mg_build<true>: bit-identical.
struct K_k_mg_build_even {
    static constexpr int MINB = 0;
    using Args = K_k_mg_build<true>::Args;
    static __device__ void run(const Args& A_, const VGrid& vg) {
        const Grid3 cg = A_.cg, fg = A_.fg;
        const uint32_t sy = (uint32_t)fg.nx, sz = (uint32_t)fg.nx * (uint32_t)fg.ny;
        const bool px = fg.per & 1u, py = fg.per & 2u, pz = fg.per & 4u;
        const bool anchors = !A_.meanfree;
        const float2* __restrict__ fdinv = reinterpret_cast<const float2*>(A_.fdinv);
        const float* __restrict__ fbx = A_.fbx;
        const float* __restrict__ fby = A_.fby;
        const float* __restrict__ fbz = A_.fbz;
        const uint32_t* __restrict__ mask = A_.mask;
        uint32_t lo, hi;
        block_range(vg, cg.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int qx, qy, qz;
            decode(cg, i, qx, qy, qz);
            const int X = 2 * qx, Y = 2 * qy, Z = 2 * qz;
            // outside neighbours: valid, and their coordinate (the child's own where invalid: any
            // in-range index, its value is never used)
            const bool xl = X > 0 || px, xh = X + 2 < fg.nx || px;
            const bool yl = Y > 0 || py, yh = Y + 2 < fg.ny || py;
            const bool zl = Z > 0 || pz, zh = Z + 2 < fg.nz || pz;
            const uint32_t xm = X > 0 ? (uint32_t)(X - 1) : (px ? (uint32_t)(fg.nx - 1) : (uint32_t)X);
            const uint32_t xp = X + 2 < fg.nx ? (uint32_t)(X + 2) : (px ? 0u : (uint32_t)X);
            const uint32_t ym = Y > 0 ? (uint32_t)(Y - 1) : (py ? (uint32_t)(fg.ny - 1) : (uint32_t)Y);
            const uint32_t yp = Y + 2 < fg.ny ? (uint32_t)(Y + 2) : (py ? 0u : (uint32_t)Y);
            const uint32_t zm = Z > 0 ? (uint32_t)(Z - 1) : (pz ? (uint32_t)(fg.nz - 1) : (uint32_t)Z);
            const uint32_t zp = Z + 2 < fg.nz ? (uint32_t)(Z + 2) : (pz ? 0u : (uint32_t)Z);
            uint32_t row[2][2];  // [dz][dy]: fine index of (X, Y + dy, Z + dz)
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++)
                    row[dz][dy] = (uint32_t)X + (uint32_t)(Y + dy) * sy + (uint32_t)(Z + dz) * sz;
            // children: 1/diag and lower faces as x pairs; upper faces of the outer children
            float2 dv[2][2], bxv[2][2], byv[2][2], bzv[2][2], byu[2], bzu[2];
            float bxu[2][2];
            uint2 mc[2][2], mym[2], myp[2], mzm[2], mzp[2];
            uint32_t mxm[2][2], mxp[2][2];
#pragma unroll
            for (int dz = 0; dz < 2; dz++)
#pragma unroll
                for (int dy = 0; dy < 2; dy++) {
                    const uint32_t r = row[dz][dy];
                    dv[dz][dy] = fdinv[r >> 1];
                    bxv[dz][dy] = *reinterpret_cast<const float2*>(fbx + r);
                    byv[dz][dy] = *reinterpret_cast<const float2*>(fby + r);
                    bzv[dz][dy] = *reinterpret_cast<const float2*>(fbz + r);
                    bxu[dz][dy] = fbx[r - (uint32_t)X + xp];
                }
#pragma unroll
            for (int d = 0; d < 2; d++) {
                // y: the row Y + 2 of plane Z + d; z: the plane Z + 2 of row Y + d
                byu[d] = *reinterpret_cast<const float2*>(fby + (uint32_t)X + yp * sy + (uint32_t)(Z + d) * sz);
                bzu[d] = *reinterpret_cast<const float2*>(fbz + (uint32_t)X + (uint32_t)(Y + d) * sy + zp * sz);
            }
            if (anchors) {
#pragma unroll
                for (int dz = 0; dz < 2; dz++)
#pragma unroll
                    for (int dy = 0; dy < 2; dy++) {
                        const uint32_t r = row[dz][dy];
                        mc[dz][dy] = *reinterpret_cast<const uint2*>(mask + r);
                        mxm[dz][dy] = mask[r - (uint32_t)X + xm];
                        mxp[dz][dy] = mask[r - (uint32_t)X + xp];
                    }
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    mym[d] = *reinterpret_cast<const uint2*>(mask + (uint32_t)X + ym * sy + (uint32_t)(Z + d) * sz);
                    myp[d] = *reinterpret_cast<const uint2*>(mask + (uint32_t)X + yp * sy + (uint32_t)(Z + d) * sz);
                    mzm[d] = *reinterpret_cast<const uint2*>(mask + (uint32_t)X + (uint32_t)(Y + d) * sy + zm * sz);
                    mzp[d] = *reinterpret_cast<const uint2*>(mask + (uint32_t)X + (uint32_t)(Y + d) * sy + zp * sz);
                }
            } else {
#pragma unroll
                for (int dz = 0; dz < 2; dz++)
#pragma unroll
                    for (int dy = 0; dy < 2; dy++) {
                        mc[dz][dy] = make_uint2(0u, 0u);
                        mxm[dz][dy] = mxp[dz][dy] = 0u;
                    }
#pragma unroll
                for (int d = 0; d < 2; d++) mym[d] = myp[d] = mzm[d] = mzp[d] = make_uint2(0u, 0u);
            }
            auto pick = [](float2 v, int s) { return s ? v.y : v.x; };
            auto pickm = [](uint2 v, int s) { return s ? v.y : v.x; };
            float c4[4] = {0.f, 0.f, 0.f, 0.f};
            float diag = 0.f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
                if (pick(dv[dz][dy], dx) == 0.0f) continue;
                // the six faces x-, x+, y-, y+, z-, z+: valid, mobility, neighbour mask, inside
                const bool ok[6] = {dx || xl, !dx || xh, dy || yl, !dy || yh, dz || zl, !dz || zh};
                const float bt[6] = {pick(bxv[dz][dy], dx),
                                     dx ? bxu[dz][dy] : bxv[dz][dy].y,
                                     pick(byv[dz][dy], dx),
                                     dy ? pick(byu[dz], dx) : pick(byv[dz][1], dx),
                                     pick(bzv[dz][dy], dx),
                                     dz ? pick(bzu[dy], dx) : pick(bzv[1][dy], dx)};
                const uint32_t mk[6] = {dx ? mc[dz][dy].x : mxm[dz][dy],
                                        dx ? mxp[dz][dy] : mc[dz][dy].y,
                                        dy ? pickm(mc[dz][0], dx) : pickm(mym[dz], dx),
                                        dy ? pickm(myp[dz], dx) : pickm(mc[dz][1], dx),
                                        dz ? pickm(mc[0][dy], dx) : pickm(mzm[dy], dx),
                                        dz ? pickm(mzp[dy], dx) : pickm(mc[1][dy], dx)};
                const bool inside[6] = {dx != 0, dx == 0, dy != 0, dy == 0, dz != 0, dz == 0};
#pragma unroll
                for (int f = 0; f < 6; f++) {
                    if (!ok[f]) continue;
                    const float beta = bt[f];
                    if (beta == 0.0f) continue;
                    if (anchors && (mk[f] & 3u) == 2u) {
                        c4[3] += beta;
                        diag += beta;
                        continue;
                    }
                    if (inside[f]) continue;
                    diag += beta;
                    if (!(f & 1)) c4[f >> 1] += beta;
                }
            }
            const float dinv = diag > 0.0f ? 1.0f / diag : 0.0f;
            A_.cbx[i] = c4[0];
            A_.cby[i] = c4[1];
            A_.cbz[i] = c4[2];
            A_.creact[i] = c4[3];
            A_.cdiag[i] = diag;
            A_.cdinv[i] = dinv;
            if (A_.c16[0]) {
                A_.c16[0][i] = __float2bfloat16(c4[0]);
                A_.c16[1][i] = __float2bfloat16(c4[1]);
                A_.c16[2][i] = __float2bfloat16(c4[2]);
                A_.c16[3][i] = __float2bfloat16(c4[3]);
                A_.c16[4][i] = __float2bfloat16(dinv);
            }
        }
    }
};

// ------------------------------------------------ row-vector forms (nx % 4 == 0)
// One thread owns four consecutive cells of a row, so cell arrays move as 16-byte loads and
// stores and every face mobility is evaluated once.  Each value is produced by the same
// FP32 expression, in the same order, as in the one-cell kernels above (bit-identical).

namespace {

// density() of a cell with liquid fraction a (solver.wgsl:40-41)
__device__ __forceinline__ float dens(const PhaseP& ph, float a) { return ph.rho_air * (1.0f - a) + ph.rho_l * a; }
// mobility of the face between cells with densities dl (lower) and dc (solver.wgsl:43-51)
__device__ __forceinline__ float mob(const PhaseP& ph, float dl, float dc) {
    if (!ph.two_phase) return 1.0f;
    float fd = 0.5f * (dl + dc);
    return ph.rho_l / fd;
}

}  // namespace

// pressure_coefficients (solver.wgsl:93-103) in one pass on four cells of a row: the three
// lower-face mobilities of each cell (and their BF16 copies), the diagonal = sum of the six face
// mobilities in the order x-, x+, y-, y+, z-, z+ (a closed face contributes +0), the anchor row
// and the active-cell reciprocal.  The upper faces are the neighbours' lower faces, evaluated
// here from the same liquid-fraction values with the same expression, so the diagonal equals the
// sum of the stored mobilities bit for bit.  Which faces are open, which cells are fluid and which
// are anchors comes from the per-cell open bytes (bit a: lower face on axis a open, i.e. the cell
// and its lower neighbour fluid; bit 3 fluid; bit 4 anchor), one byte per cell instead of the
// four-byte gauge mask of the cell and its six neighbours.
struct K_k_coef4 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        PhaseP ph;
        const uint8_t* open8;
        const float* alpha;
        float* bx;
        float* by;
        float* bz;
        float* diag;
        float* dinv;
        __hip_bfloat16* bx16;
        __hip_bfloat16* by16;
        __hip_bfloat16* bz16;
        __hip_bfloat16* dinv16;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        const Grid3 g = A_.g;
        const PhaseP ph = A_.ph;
        const uint8_t* __restrict__ open8 = A_.open8;
        const float* __restrict__ alpha = A_.alpha;
        uint32_t lo, hi;
        block_range(vg, g.n >> 2, lo, hi);
        for (uint32_t t = lo + threadIdx.x; t < hi; t += MHDP_BLOCK) {
            const Row4 r = row4(g, t);
            const uint32_t ow = *reinterpret_cast<const uint32_t*>(open8 + r.i0);
            const uint32_t oyu = *reinterpret_cast<const uint32_t*>(open8 + r.yp);
            const uint32_t ozu = *reinterpret_cast<const uint32_t*>(open8 + r.zp);
            const uint32_t oxu = open8[r.xp];
            auto byte = [](uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; };
            float a[4] = {0.f, 0.f, 0.f, 0.f}, ay[4] = {0.f, 0.f, 0.f, 0.f}, az[4] = {0.f, 0.f, 0.f, 0.f};
            float ayu[4] = {0.f, 0.f, 0.f, 0.f}, azu[4] = {0.f, 0.f, 0.f, 0.f}, ax = 0.f, axu = 0.f;
            if (alpha) {
                ld4f(alpha, r.i0, a);
                ld4f(alpha, r.ym, ay);
                ld4f(alpha, r.zm, az);
                ld4f(alpha, r.yp, ayu);
                ld4f(alpha, r.zp, azu);
                ax = alpha[r.xm];
                axu = alpha[r.xp];
            }
            float ox[4], oy[4], oz[4], d[4], di[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t o = byte(ow, k);
                const float dc = dens(ph, a[k]);
                // lower faces (mobility(), solver.wgsl:43-51)
                ox[k] = (o & 1u) ? mob(ph, dens(ph, k > 0 ? a[k - 1] : ax), dc) : 0.0f;
                oy[k] = (o & 2u) ? mob(ph, dens(ph, ay[k]), dc) : 0.0f;
                oz[k] = (o & 4u) ? mob(ph, dens(ph, az[k]), dc) : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t o = byte(ow, k);
                const bool fc = (o & 8u) != 0u;
                const float dc = dens(ph, a[k]);
                // upper faces: the lower faces of the +x, +y, +z neighbours
                const float xu = k < 3 ? ox[k + 1] : ((r.xp_ok && (oxu & 1u)) ? mob(ph, dc, dens(ph, axu)) : 0.0f);
                const float yu = (r.yp_ok && (byte(oyu, k) & 2u)) ? mob(ph, dc, dens(ph, ayu[k])) : 0.0f;
                const float zu = (r.zp_ok && (byte(ozu, k) & 4u)) ? mob(ph, dc, dens(ph, azu[k])) : 0.0f;
                // diagonal() (solver.wgsl:60-69)
                float v = 0.0f;
                if (fc) {
                    v += ox[k];
                    v += xu;
                    v += oy[k];
                    v += yu;
                    v += oz[k];
                    v += zu;
                    if (!ph.meanfree && (o & 16u)) v = 1.0f;
                    else v = fmaxf(v, 1e-30f);
                }
                d[k] = v;
                const bool active = fc && (ph.meanfree || !(o & 16u));
                di[k] = active ? 1.0f / v : 0.0f;
            }
            st4f(A_.bx, r.i0, ox);
            st4f(A_.by, r.i0, oy);
            st4f(A_.bz, r.i0, oz);
            st4f(A_.diag, r.i0, d);
            st4f(A_.dinv, r.i0, di);
            if (A_.bx16) {
                st4h(A_.bx16, r.i0, ox);
                st4h(A_.by16, r.i0, oy);
                st4h(A_.bz16, r.i0, oz);
                st4h(A_.dinv16, r.i0, di);
            }
        }
    }
};

// rhs() (solver.wgsl:104-107) on the four cells of row r, out[k] = -h^2 div: faces of the x plane
// are read as five scalars (its rows hold nx+1 faces), the y and z planes as 16-byte rows.  A face
// counts where it is open (its mobility is nonzero): bit a of the open byte of the cell above it
// on axis a; bit 3 marks a fluid cell.
template <class T>
__device__ __forceinline__ void rhs_row4(const Grid3& g, float h, const Row4& r, const uint8_t* __restrict__ open8,
                                         const float* __restrict__ U, const float* __restrict__ V,
                                         const float* __restrict__ W, T (&out)[4]) {
    const uint32_t ow = *reinterpret_cast<const uint32_t*>(open8 + r.i0);
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = (T)0;
    if ((ow & 0x08080808u) == 0u) return;
    const uint32_t oyu = r.yp_ok ? *reinterpret_cast<const uint32_t*>(open8 + r.yp) : 0u;
    const uint32_t ozu = r.zp_ok ? *reinterpret_cast<const uint32_t*>(open8 + r.zp) : 0u;
    const uint32_t oxu = r.xp_ok ? (uint32_t)open8[r.xp] : 0u;
    auto byte = [](uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; };
    // x plane: faces x0..x0+4 of this row (the last wraps to face 0 when periodic)
    const uint32_t ux = (uint32_t)(g.nx + 1) * ((uint32_t)r.y + (uint32_t)g.ny * (uint32_t)r.z);
    float uf[5];
#pragma unroll
    for (int k = 0; k < 4; k++) uf[k] = U[ux + (uint32_t)r.x0 + (uint32_t)k];
    const bool wrapx = r.x0 + 4 >= g.nx && (g.per & 1u);
    uf[4] = U[ux + (wrapx ? 0u : (uint32_t)r.x0 + 4u)];
    // y plane rows y and y+1 (y+1 wraps to 0 when periodic)
    const uint32_t Py = (uint32_t)g.ny + 1u;
    const uint32_t vy = (uint32_t)r.x0 + (uint32_t)g.nx * ((uint32_t)r.y + Py * (uint32_t)r.z);
    const int yu_i = r.y + 1 < g.ny ? r.y + 1 : ((g.per & 2u) ? 0 : r.y + 1);
    const uint32_t vyu = (uint32_t)r.x0 + (uint32_t)g.nx * ((uint32_t)yu_i + Py * (uint32_t)r.z);
    float vl[4], vu[4];
    ld4f(V, vy, vl);
    ld4f(V, vyu, vu);
    // z plane rows z and z+1
    const uint32_t sz = (uint32_t)g.nx * (uint32_t)g.ny;
    const int zu_i = r.z + 1 < g.nz ? r.z + 1 : ((g.per & 4u) ? 0 : r.z + 1);
    const uint32_t wz = (uint32_t)r.x0 + (uint32_t)g.nx * (uint32_t)r.y;
    float wl[4], wu[4];
    ld4f(W, wz + sz * (uint32_t)r.z, wl);
    ld4f(W, wz + sz * (uint32_t)zu_i, wu);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t o = byte(ow, k);
        if (!(o & 8u)) continue;
        const bool xu_open = k < 3 ? (byte(ow, k + 1) & 1u) != 0u : (oxu & 1u) != 0u;
        T sum = (T)0;
        {
            const float up = xu_open ? uf[k + 1] : 0.0f;
            const float dn = (o & 1u) ? uf[k] : 0.0f;
            sum += (T)up - (T)dn;
        }
        {
            const float up = (byte(oyu, k) & 2u) ? vu[k] : 0.0f;
            const float dn = (o & 2u) ? vl[k] : 0.0f;
            sum += (T)up - (T)dn;
        }
        {
            const float up = (byte(ozu, k) & 4u) ? wu[k] : 0.0f;
            const float dn = (o & 4u) ? wl[k] : 0.0f;
            sum += (T)up - (T)dn;
        }
        T div = sum / (T)h;
        out[k] = -(T)h * (T)h * div;
    }
}

template <class T>
struct K_k_rhs4 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float h;
        const uint8_t* open8;
        const float* ustar;
        uint64_t off1;
        uint64_t off2;
        T* b;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        const Grid3 g = A_.g;
        uint32_t lo, hi;
        block_range(vg, g.n >> 2, lo, hi);
        const float* __restrict__ U = A_.ustar;
        for (uint32_t t = lo + threadIdx.x; t < hi; t += MHDP_BLOCK) {
            const Row4 r = row4(g, t);
            T out[4];
            rhs_row4<T>(g, A_.h, r, A_.open8, U, U + A_.off1, U + A_.off2, out);
#pragma unroll
            for (int k = 0; k < 4; k++) A_.b[r.i0 + k] = out[k];
        }
    }
};

// The FP64 right-hand side and the first refinement step's start from a zero potential in one
// pass (K_k_outer_start4 with zero_x: the FP32 inner right-hand side r32 = b on active cells, the
// inner x and both search directions zeroed, the partials of FIN_OUTER).  The sums are the
// fused multiply-adds that pass performs.
struct K_k_rhs_start4 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float h;
        const uint8_t* open8;
        const float* ustar;
        uint64_t off1;
        uint64_t off2;
        double* b;
        int meanfree;  // active cells: fluid and, in the pinned gauge, not an anchor
        float* r32;
        float* x32;
        float* p0;
        float* p1;
        double* part;
        Scal* sc;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        const Grid3 g = A_.g;
        uint32_t lo, hi;
        block_range(vg, g.n >> 2, lo, hi);
        const float* __restrict__ U = A_.ustar;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const float zero4[4] = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t t = lo + threadIdx.x; t < hi; t += MHDP_BLOCK) {
            const Row4 r = row4(g, t);
            double bv[4];
            rhs_row4<double>(g, A_.h, r, A_.open8, U, U + A_.off1, U + A_.off2, bv);
            st4d(A_.b, r.i0, bv);
            const uint32_t ow = *reinterpret_cast<const uint32_t*>(A_.open8 + r.i0);
            float rv[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t o = (ow >> (8 * k)) & 0xFFu;
                const bool active = (o & 8u) && (A_.meanfree || !(o & 16u));
                rv[k] = active ? (float)bv[k] : 0.0f;
                acc[0] = __fma_rn((double)rv[k], (double)rv[k], acc[0]);
                acc[1] = __fma_rn(bv[k], bv[k], acc[1]);
                acc[2] = __fma_rn(bv[k], bv[k], acc[2]);
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

// correct() (solver.wgsl:245-250) on four cells of a row: each thread writes the lower faces
// of its cells in all three planes, and the planes' last faces (x = nx, y = ny, z = nz) where
// its cells border them (the periodic copy of face 0, or 0 when closed).
template <class PT>
__device__ __forceinline__ float corrected(float uf, float beta, PT pc, PT pl, float h) {
    if (sizeof(PT) == 4) return uf - (float)beta * ((float)pc - (float)pl) / h;
    double o = (double)uf - (double)beta * ((double)pc - (double)pl) / (double)h;
    return (float)o;
}
__device__ __forceinline__ void ld4p(const float* __restrict__ p, uint32_t i, float (&v)[4]) { ld4f(p, i, v); }
__device__ __forceinline__ void ld4p(const double* __restrict__ p, uint32_t i, double (&v)[4]) {
    const double2 a = *reinterpret_cast<const double2*>(p + i), b = *reinterpret_cast<const double2*>(p + i + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void st4p(float* __restrict__ p, uint32_t i, const float (&v)[4]) { st4f(p, i, v); }
__device__ __forceinline__ void st4p(double* __restrict__ p, uint32_t i, const double (&v)[4]) {
    *reinterpret_cast<double2*>(p + i) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + i + 2) = make_double2(v[2], v[3]);
}
template <class PT>
struct K_k_correct4 {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float h;
        const float* bx;
        const float* by;
        const float* bz;
        const PT* phi;
        const float* ustar;
        float* uout;
        uint32_t f0;
        uint32_t f1;
        PT* phi_copy;  // when set: the potential is copied here in the same pass
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        Grid3 g = A_.g;
        float h = A_.h;
        const float* __restrict__ bx = A_.bx;
        const float* __restrict__ by = A_.by;
        const float* __restrict__ bz = A_.bz;
        const PT* __restrict__ phi = A_.phi;
        const float* __restrict__ ustar = A_.ustar;
        float* __restrict__ uout = A_.uout;
        uint32_t f0 = A_.f0;
        uint32_t f1 = A_.f1;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, g.n >> 2, lo, hi);
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const float* __restrict__ U = ustar;
        const float* __restrict__ V = ustar + f0;
        const float* __restrict__ W = ustar + f0 + f1;
        float* __restrict__ UO = uout;
        float* __restrict__ VO = uout + f0;
        float* __restrict__ WO = uout + f0 + f1;
        for (uint32_t t = lo + threadIdx.x; t < hi; t += MHDP_BLOCK) {
            const Row4 r = row4(g, t);
            PT pc[4], py[4], pz[4];
            ld4p(phi, r.i0, pc);
            if (A_.phi_copy) st4p(A_.phi_copy, r.i0, pc);
            ld4p(phi, r.ym, py);
            ld4p(phi, r.zm, pz);
            const PT px = phi[r.xm];
            float xl[4], yl[4], zl[4];
            ld4f(bx, r.i0, xl);
            ld4f(by, r.i0, yl);
            ld4f(bz, r.i0, zl);
            // x plane
            const uint32_t ux = (uint32_t)(g.nx + 1) * ((uint32_t)r.y + (uint32_t)g.ny * (uint32_t)r.z) + (uint32_t)r.x0;
    #pragma unroll
            for (int k = 0; k < 4; k++) {
                float o = 0.0f;
                if (xl[k] > 0.0f) o = corrected<PT>(U[ux + k], xl[k], pc[k], k > 0 ? pc[k - 1] : px, h);
                UO[ux + k] = o;
            }
            if (r.x0 + 4 == g.nx) {  // face x = nx of this row
                float o = 0.0f;
                if (g.per & 1u) {
                    const uint32_t c0 = r.i0 + 4u - (uint32_t)g.nx;  // cell x = 0 of the row
                    const float b0 = bx[c0];
                    if (b0 > 0.0f) o = corrected<PT>(U[ux + 4u - (uint32_t)g.nx], b0, phi[c0], pc[3], h);
                }
                UO[ux + 4] = o;
            }
            // y plane
            const uint32_t Py = (uint32_t)g.ny + 1u;
            const uint32_t vy = (uint32_t)r.x0 + sy * ((uint32_t)r.y + Py * (uint32_t)r.z);
            {
                float v[4], o[4];
                ld4f(V, vy, v);
    #pragma unroll
                for (int k = 0; k < 4; k++) o[k] = yl[k] > 0.0f ? corrected<PT>(v[k], yl[k], pc[k], py[k], h) : 0.0f;
                st4f(VO, vy, o);
            }
            if (r.y + 1 == g.ny) {  // faces y = ny
                float o[4] = {0.f, 0.f, 0.f, 0.f};
                if (g.per & 2u) {
                    const uint32_t c0 = r.i0 - (uint32_t)(g.ny - 1) * sy;  // row y = 0
                    const uint32_t v0 = (uint32_t)r.x0 + sy * (Py * (uint32_t)r.z);
                    float b0[4], v[4];
                    ld4f(by, c0, b0);
                    ld4f(V, v0, v);
                    PT p0[4];
                    ld4p(phi, c0, p0);
    #pragma unroll
                    for (int k = 0; k < 4; k++) o[k] = b0[k] > 0.0f ? corrected<PT>(v[k], b0[k], p0[k], pc[k], h) : 0.0f;
                }
                st4f(VO, vy + sy, o);
            }
            // z plane
            {
                float w[4], o[4];
                ld4f(W, r.i0, w);
    #pragma unroll
                for (int k = 0; k < 4; k++) o[k] = zl[k] > 0.0f ? corrected<PT>(w[k], zl[k], pc[k], pz[k], h) : 0.0f;
                st4f(WO, r.i0, o);
            }
            if (r.z + 1 == g.nz) {  // faces z = nz
                float o[4] = {0.f, 0.f, 0.f, 0.f};
                if (g.per & 4u) {
                    const uint32_t c0 = r.i0 - (uint32_t)(g.nz - 1) * sz;  // plane z = 0
                    float b0[4], w[4];
                    ld4f(bz, c0, b0);
                    ld4f(W, c0, w);
                    PT p0[4];
                    ld4p(phi, c0, p0);
    #pragma unroll
                    for (int k = 0; k < 4; k++) o[k] = b0[k] > 0.0f ? corrected<PT>(w[k], b0[k], p0[k], pc[k], h) : 0.0f;
                }
                st4f(WO, r.i0 + sz, o);
            }
        }
    }
};

struct K_k_to_bf16 {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const float* src;
        __hip_bfloat16* dst;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        uint32_t n = A_.n;
        const float* __restrict__ src = A_.src;
        __hip_bfloat16* __restrict__ dst = A_.dst;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) dst[i] = __float2bfloat16(src[i]);
    }
};

struct K_k_from_bf16 {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const __hip_bfloat16* src;
        float* dst;
    };
    static __device__ void run(const Args& A_, const VGrid& vg) {
        uint32_t n = A_.n;
        const __hip_bfloat16* __restrict__ src = A_.src;
        float* __restrict__ dst = A_.dst;
        (void)vg;
        uint32_t lo, hi;
        block_range(vg, n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) dst[i] = __bfloat162float(src[i]);
    }
};

// ---------------------------------------------------------------- warm start
// old_potential() (solver.wgsl:164-167): the previous potential relative to the cell's component
// anchor on fluid cells, 0 elsewhere.
__device__ __forceinline__ double old_potential(const uint32_t* __restrict__ mask, const double* __restrict__ prev,
                                                uint32_t i) {
    const uint32_t m = mask[i];
    return m ? prev[i] - prev[m >> 2] : 0.0;
}

// pressure_warm (solver.wgsl:170-176): u* - beta (phi_old(c) - phi_old(c - e_a)) / h on every
// open face; closed faces keep u*.
struct K_warm {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float h;
        const uint32_t* mask;
        const double* prev;
        const float* bx;
        const float* by;
        const float* bz;
        const float* ustar;
        float* uout;
        uint32_t f0, f1, f2;
    };
    static __device__ void run(const Args& A, const VGrid& vg) {
        const Grid3 g = A.g;
        const uint32_t total = A.f0 + A.f1 + A.f2;
        uint32_t lo, hi;
        block_range(vg, total, lo, hi);
        const float* B[3] = {A.bx, A.by, A.bz};
        const int d[3] = {g.nx, g.ny, g.nz};
        const uint32_t st[3] = {1u, (uint32_t)g.nx, (uint32_t)g.nx * (uint32_t)g.ny};
        for (uint32_t f = lo + threadIdx.x; f < hi; f += MHDP_BLOCK) {
            int a;
            uint32_t k;
            if (f < A.f0) { a = 0; k = f; }
            else if (f < A.f0 + A.f1) { a = 1; k = f - A.f0; }
            else { a = 2; k = f - A.f0 - A.f1; }
            const uint32_t Px = (uint32_t)g.nx + (a == 0), Py = (uint32_t)g.ny + (a == 1);
            int q[3];
            q[0] = (int)(k % Px);
            q[1] = (int)((k / Px) % Py);
            q[2] = (int)(k / (Px * Py));
            float out = A.ustar[f];
            bool ok = true;
            if (q[a] >= d[a]) {
                if (g.per & (1u << a)) q[a] = 0;
                else ok = false;
            }
            if (ok) {
                const uint32_t ci = (uint32_t)q[0] + (uint32_t)g.nx * ((uint32_t)q[1] + (uint32_t)g.ny * (uint32_t)q[2]);
                const float beta = B[a][ci];
                if (beta > 0.0f) {
                    const uint32_t li = q[a] > 0 ? ci - st[a] : ci + (uint32_t)(d[a] - 1) * st[a];
                    const double dp = old_potential(A.mask, A.prev, ci) - old_potential(A.mask, A.prev, li);
                    out = (float)((double)out - (double)beta * dp / (double)A.h);
                }
            }
            A.uout[f] = out;
        }
    }
};

// The previous potential as a starting value: phi = phi_old (0 off the fluid).
struct K_old_potential {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const uint32_t* mask;
        const double* prev;
        double* phi;
    };
    static __device__ void run(const Args& A, const VGrid& vg) {
        uint32_t lo, hi;
        block_range(vg, A.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) A.phi[i] = old_potential(A.mask, A.prev, i);
    }
};

// pressure_commit (solver.wgsl:181-183): phi += phi_old on fluid cells.
struct K_commit {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const uint32_t* mask;
        const double* prev;
        double* phi;
    };
    static __device__ void run(const Args& A, const VGrid& vg) {
        uint32_t lo, hi;
        block_range(vg, A.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK)
            if (A.mask[i]) A.phi[i] = A.phi[i] + old_potential(A.mask, A.prev, i);
    }
};

// ---------------------------------------------------------------- display packet
// fields() (solver.wgsl:256-276): the 16-float FluidCell packet per cell from the projected faces
// and the potential, in the reference's FP32 operation order, and the diagnostics() totals
// (max speed, max |div|, sum div^2, sum kinetic energy, max |p|, sum div, max |force|,
// non-finite cells, max |face velocity|, sum melt) as FP64 per-block partials.
__device__ __forceinline__ bool fields_is_max(int k) { return k == 0 || k == 1 || k == 4 || k == 6 || k == 8; }

struct K_fields {
    static constexpr int MINB = 0;
    struct Args {
        Grid3 g;
        float h, rho_l, rho_air, dt;
        float acc[3];
        int two_phase, mushy, meanfree;
        const uint32_t* mask;
        const uint8_t* open8;  // per cell: bit a = its lower face on axis a is open
        const float* faces;
        uint32_t f0, f1;
        const double* phi;
        const float* em;
        float* cells;
        double* part;
    };
    static __device__ void run(const Args& A, const VGrid& vg) {
        const Grid3 g = A.g;
        const int d[3] = {g.nx, g.ny, g.nz};
        uint32_t lo, hi;
        block_range(vg, g.n, lo, hi);
        double tot[kFieldTotals];
#pragma unroll
        for (int k = 0; k < kFieldTotals; k++) tot[k] = 0.0;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            float* cell = A.cells + 16u * (size_t)i;
            const uint32_t m = A.mask[i];
            int x, y, z;
            decode(g, i, x, y, z);
            const int pos[3] = {x, y, z};
            // face(pos, a, false) and face(pos + e_a, a, false) (solver.wgsl:33-36): open from the
            // open bytes of the cell and of its upper neighbour (periodic: wrapped; closed: none);
            // every load below is issued from a valid index before any value is used
            uint32_t ob[4];
            float fv[6];
            bool up_ok[3];
            ob[0] = A.open8[i];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const uint32_t Px = (uint32_t)g.nx + (a == 0), Py = (uint32_t)g.ny + (a == 1);
                const uint32_t off = a == 0 ? 0u : (a == 1 ? A.f0 : A.f0 + A.f1);
                int u[3] = {x, y, z};
                u[a] += 1;
                const bool wrap = u[a] == d[a] && (g.per & (1u << a));
                up_ok[a] = u[a] < d[a] || wrap;
                if (wrap) u[a] = 0;
                const uint32_t uc = up_ok[a] ? (uint32_t)u[0] + (uint32_t)g.nx * ((uint32_t)u[1] + (uint32_t)g.ny * (uint32_t)u[2]) : i;
                ob[1 + a] = A.open8[uc];
                fv[2 * a] = A.faces[off + (uint32_t)x + Px * ((uint32_t)y + Py * (uint32_t)z)];
                fv[2 * a + 1] = A.faces[off + (uint32_t)u[0] + Px * ((uint32_t)u[1] + Py * (uint32_t)u[2])];
            }
            const float o1 = cell[13], o2 = cell[14], o3 = cell[15];
            const double phi_i = A.phi[i], phi_a = A.meanfree ? A.phi[m >> 2] : 0.0;
            float force[3] = {0.0f, 0.0f, 0.0f};
            if (A.em)
                for (int a = 0; a < 3; a++) force[a] = A.em[16u * (size_t)i + 8u + (uint32_t)a];
#pragma unroll
            for (int k = 0; k < 4; k++) asm volatile("" : "+v"(ob[k]));
#pragma unroll
            for (int k = 0; k < 6; k++) asm volatile("" : "+v"(fv[k]));
            float out[16];
#pragma unroll
            for (int k = 0; k < 16; k++) out[k] = 0.0f;
            if (A.two_phase) { out[13] = o1; out[14] = o2; out[15] = o3; }
            if (m) {
                float v[3], lo_[3], hi_[3], fe = 0.0f, mf = 0.0f, sum = 0.0f;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    lo_[a] = ((ob[0] >> a) & 1u) ? fv[2 * a] : 0.0f;
                    hi_[a] = (up_ok[a] && ((ob[1 + a] >> a) & 1u)) ? fv[2 * a + 1] : 0.0f;
                    v[a] = 0.5f * (lo_[a] + hi_[a]);
                    fe += 0.5f * (lo_[a] * lo_[a] + hi_[a] * hi_[a]);
                    mf = fmaxf(mf, fmaxf(fabsf(lo_[a]), fabsf(hi_[a])));
                }
#pragma unroll
                for (int a = 0; a < 3; a++) sum += hi_[a] - lo_[a];
                const float div = sum / A.h;
                float melt = A.mushy ? o1 : 1.0f;
                if (A.two_phase) melt *= o2;
                const float speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                const float pv = A.meanfree ? (float)phi_i - (float)phi_a : (float)phi_i;
                const float pressure = pv * A.rho_l / A.dt;
#pragma unroll
                for (int a = 0; a < 3; a++) out[a] = v[a];
                out[3] = speed;
                out[4] = pressure;
                out[5] = div;
#pragma unroll
                for (int a = 0; a < 3; a++) out[8 + a] = A.rho_l * A.acc[a] + force[a];
                out[12] = 1.0f;
                out[13] = A.mushy ? o1 : 1.0f;
                out[14] = o2;
                out[15] = o3;
                bool bad = false;
#pragma unroll
                for (int a = 0; a < 3; a++) bad |= v[a] != v[a] || force[a] != force[a] || fabsf(force[a]) > 1e30f;
                bad |= speed > 1e25f || !(fabsf(pressure) < 1e30f) || !(fabsf(div) < 1e30f);
                const float rho = A.two_phase ? A.rho_air * (1.0f - o2) + A.rho_l * o2 : A.rho_l;
                const float ke = 0.5f * rho * fe * A.h * A.h * A.h;
                const float fl = sqrtf(force[0] * force[0] + force[1] * force[1] + force[2] * force[2]);
                tot[0] = fmax(tot[0], (double)speed);
                tot[1] = fmax(tot[1], (double)fabsf(div));
                tot[2] += (double)(div * div);
                tot[3] += (double)ke;
                tot[4] = fmax(tot[4], (double)fabsf(pressure));
                tot[5] += (double)div;
                tot[6] = fmax(tot[6], (double)fl);
                tot[7] += bad ? 1.0 : 0.0;
                tot[8] = fmax(tot[8], (double)mf);
                tot[9] += (double)melt;
            }
            float4* c4 = reinterpret_cast<float4*>(cell);
#pragma unroll
            for (int k = 0; k < 4; k++) c4[k] = make_float4(out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3]);
        }
        // block totals (fixed order), one partial record per block
        __shared__ double red[MHDP_BLOCK / 64][kFieldTotals];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < kFieldTotals; k++) {
            double v = tot[k];
            for (int o = 32; o > 0; o >>= 1) {
                const double w = __shfl_down(v, o, 64);
                v = fields_is_max(k) ? fmax(v, w) : v + w;
            }
            if (lane == 0) red[wid][k] = v;
        }
        __syncthreads();
        if (threadIdx.x < kFieldTotals) {
            const int k = threadIdx.x;
            double v = red[0][k];
            for (int w = 1; w < MHDP_BLOCK / 64; w++) v = fields_is_max(k) ? fmax(v, red[w][k]) : v + red[w][k];
            A.part[(size_t)kFieldTotals * vg.b + (uint32_t)k] = v;
        }
    }
};

// diagnostics() (solver.wgsl:284-291): fold the per-block partials - thread t takes records
// t, t + 256, ... in order, then a fixed tree over the threads (deterministic).
struct K_fields_total {
    static constexpr int MINB = 0;
    struct Args {
        const double* part;
        uint32_t blocks;
        double* out;
    };
    static __device__ void run(const Args& A, const VGrid&) {
        __shared__ double red[kFieldTotals][MHDP_BLOCK];
        double v[kFieldTotals];
#pragma unroll
        for (int k = 0; k < kFieldTotals; k++) v[k] = 0.0;
        for (uint32_t b = threadIdx.x; b < A.blocks; b += MHDP_BLOCK)
#pragma unroll
            for (int k = 0; k < kFieldTotals; k++) {
                const double w = A.part[(size_t)kFieldTotals * b + (uint32_t)k];
                v[k] = fields_is_max(k) ? fmax(v[k], w) : v[k] + w;
            }
#pragma unroll
        for (int k = 0; k < kFieldTotals; k++) red[k][threadIdx.x] = v[k];
        __syncthreads();
        for (int s = MHDP_BLOCK / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s)
#pragma unroll
                for (int k = 0; k < kFieldTotals; k++) {
                    const double a = red[k][threadIdx.x], b = red[k][threadIdx.x + s];
                    red[k][threadIdx.x] = fields_is_max(k) ? fmax(a, b) : a + b;
                }
            __syncthreads();
        }
        if (threadIdx.x < kFieldTotals) A.out[threadIdx.x] = red[threadIdx.x][0];
    }
};

// ---------------------------------------------------------------- launchers

static bool row_vector_ok(const Grid3& g) { return (g.nx & 3) == 0 && std::getenv("MHDP_SCALAR_SETUP") == nullptr; }

bool launch_coefficients(hipStream_t s, const Grid3& g, const PhaseP& ph, const uint32_t* mask, const uint8_t* open8,
                         const float* alpha,
                         float* bx, float* by, float* bz, float* diag, float* dinv, void* bx16, void* by16,
                         void* bz16, void* dinv16) {
    if (row_vector_ok(g) && open8) {
        const dim3 G(grid_blocks(g.n >> 2)), B(MHDP_BLOCK);
        const bool h = bx16 && by16 && bz16 && dinv16;
        using T = __hip_bfloat16;
        go<K_k_coef4>(s, G.x, {g, ph, open8, alpha, bx, by, bz, diag, dinv, h ? (T*)bx16 : (T*)nullptr,
                               h ? (T*)by16 : (T*)nullptr, h ? (T*)bz16 : (T*)nullptr, h ? (T*)dinv16 : (T*)nullptr});
        return h;
    }
    go<K_k_coefficients>(s, grid_blocks(g.n), {g, ph, mask, alpha, bx, by, bz, diag, dinv});
    return false;
}
void launch_rhs_f32(hipStream_t s, const Grid3& g, float h, const uint32_t* mask, const uint8_t* open8, const float* bx,
                    const float* by, const float* bz, const float* ustar, float* b) {
    uint64_t f0 = (uint64_t)(g.nx + 1) * g.ny * g.nz, f1 = (uint64_t)g.nx * (g.ny + 1) * g.nz;
    if (row_vector_ok(g) && open8)
        go<K_k_rhs4<float>>(s, grid_blocks(g.n >> 2), {g, h, open8, ustar, f0, f0 + f1, b});
    else
        go<K_k_rhs<float>>(s, grid_blocks(g.n), {g, h, mask, bx, by, bz, ustar, f0, f0 + f1, b});
}
bool launch_rhs_start(hipStream_t s, const Grid3& g, float h, const uint8_t* open8, const float* ustar, double* b,
                      bool meanfree, float* r32, float* x32, float* p0, float* p1, double* part, Scal* sc) {
    if (!row_vector_ok(g) || !open8) return false;
    uint64_t f0 = (uint64_t)(g.nx + 1) * g.ny * g.nz, f1 = (uint64_t)g.nx * (g.ny + 1) * g.nz;
    go<K_k_rhs_start4>(s, grid_blocks(g.n >> 2),
                       {g, h, open8, ustar, f0, f0 + f1, b, meanfree ? 1 : 0, r32, x32, p0, p1, part, sc});
    return true;
}
void launch_rhs_f64(hipStream_t s, const Grid3& g, float h, const uint32_t* mask, const uint8_t* open8, const float* bx,
                    const float* by, const float* bz, const float* ustar, double* b) {
    uint64_t f0 = (uint64_t)(g.nx + 1) * g.ny * g.nz, f1 = (uint64_t)g.nx * (g.ny + 1) * g.nz;
    if (row_vector_ok(g) && open8)
        go<K_k_rhs4<double>>(s, grid_blocks(g.n >> 2), {g, h, open8, ustar, f0, f0 + f1, b});
    else
        go<K_k_rhs<double>>(s, grid_blocks(g.n), {g, h, mask, bx, by, bz, ustar, f0, f0 + f1, b});
}
static void face_counts(const Grid3& g, uint32_t& f0, uint32_t& f1, uint32_t& f2) {
    f0 = (uint32_t)((g.nx + 1) * (uint64_t)g.ny * g.nz);
    f1 = (uint32_t)(g.nx * (uint64_t)(g.ny + 1) * g.nz);
    f2 = (uint32_t)(g.nx * (uint64_t)g.ny * (g.nz + 1));
}
void launch_correct_f32(hipStream_t s, const Grid3& g, float h, const float* bx, const float* by, const float* bz,
                        const float* phi, const float* ustar, float* uout) {
    uint32_t f0, f1, f2;
    face_counts(g, f0, f1, f2);
    if (row_vector_ok(g))
        go<K_k_correct4<float>>(s, grid_blocks(g.n >> 2), {g, h, bx, by, bz, phi, ustar, uout, f0, f1, nullptr});
    else
        go<K_k_correct<float>>(s, grid_blocks(f0 + f1 + f2), {g, h, bx, by, bz, phi, ustar, uout, f0, f1, f2});
}
bool launch_correct_f64(hipStream_t s, const Grid3& g, float h, const float* bx, const float* by, const float* bz,
                        const double* phi, const float* ustar, float* uout, double* phi_copy) {
    uint32_t f0, f1, f2;
    face_counts(g, f0, f1, f2);
    if (row_vector_ok(g)) {
        go<K_k_correct4<double>>(s, grid_blocks(g.n >> 2), {g, h, bx, by, bz, phi, ustar, uout, f0, f1, phi_copy});
        return phi_copy != nullptr;
    }
    go<K_k_correct<double>>(s, grid_blocks(f0 + f1 + f2), {g, h, bx, by, bz, phi, ustar, uout, f0, f1, f2});
    return false;
}
void launch_mg_build(hipStream_t s, const Grid3& cg, const Grid3& fg, bool fine, int meanfree, const uint32_t* mask,
                     const float* fdinv, const float* fbx, const float* fby, const float* fbz, const float* freact,
                     float* cbx, float* cby, float* cbz, float* creact, float* cdiag, float* cdinv,
                     void* const c16[5]) {
    __hip_bfloat16* o16[5];
    for (int k = 0; k < 5; k++) o16[k] = c16 ? (__hip_bfloat16*)c16[k] : nullptr;
    const bool even = fg.nx % 2 == 0 && fg.ny % 2 == 0 && fg.nz % 2 == 0 && fg.nx >= 4 && fg.ny >= 4 && fg.nz >= 4;
    if (fine && even && std::getenv("MHDP_SCALAR_SETUP") == nullptr)
        go<K_k_mg_build_even>(s, grid_blocks(cg.n), {cg, fg, meanfree, mask, fdinv, fbx, fby, fbz, freact, cbx, cby, cbz, creact, cdiag,
                                                     cdinv, {o16[0], o16[1], o16[2], o16[3], o16[4]}});
    else if (fine)
        go<K_k_mg_build<true>>(s, grid_blocks(cg.n), {cg, fg, meanfree, mask, fdinv, fbx, fby, fbz, freact, cbx, cby, cbz, creact, cdiag, cdinv,
                                                      {o16[0], o16[1], o16[2], o16[3], o16[4]}});
    else
        go<K_k_mg_build<false>>(s, grid_blocks(cg.n), {cg, fg, meanfree, mask, fdinv, fbx, fby, fbz, freact, cbx, cby, cbz, creact, cdiag, cdinv,
                                                       {o16[0], o16[1], o16[2], o16[3], o16[4]}});
}
void launch_to_bf16(hipStream_t s, uint32_t n, const float* src, void* dst) {
    go<K_k_to_bf16>(s, grid_blocks(n), {n, src, (__hip_bfloat16*)dst});
}
void launch_from_bf16(hipStream_t s, uint32_t n, const void* src, float* dst) {
    go<K_k_from_bf16>(s, grid_blocks(n), {n, (const __hip_bfloat16*)src, dst});
}

void launch_warm(hipStream_t s, const Grid3& g, float h, const uint32_t* mask, const double* prev, const float* bx,
                 const float* by, const float* bz, const float* ustar, float* uout) {
    uint32_t f0, f1, f2;
    face_counts(g, f0, f1, f2);
    go<K_warm>(s, grid_blocks(f0 + f1 + f2), {g, h, mask, prev, bx, by, bz, ustar, uout, f0, f1, f2});
}
void launch_commit(hipStream_t s, uint32_t n, const uint32_t* mask, const double* prev, double* phi) {
    go<K_commit>(s, grid_blocks(n), {n, mask, prev, phi});
}
void launch_old_potential(hipStream_t s, uint32_t n, const uint32_t* mask, const double* prev, double* phi) {
    go<K_old_potential>(s, grid_blocks(n), {n, mask, prev, phi});
}

void launch_fields(hipStream_t s, const Grid3& g, float h, float rho_l, float rho_air, int two_phase, float dt,
                   const float acc[3], int mushy, int meanfree, const uint32_t* mask, const uint8_t* open8,
                   const float* faces, const double* phi, const float* em, float* cells, double* part,
                   double* totals) {
    uint32_t f0, f1, f2;
    face_counts(g, f0, f1, f2);
    const uint32_t blocks = grid_blocks(g.n);
    K_fields::Args a{g, h, rho_l, rho_air, dt, {acc[0], acc[1], acc[2]}, two_phase, mushy, meanfree, mask, open8, faces,
                     f0, f1, phi, em, cells, part};
    go<K_fields>(s, blocks, a);
    go<K_fields_total>(s, 1, {part, blocks, totals});
}
