// Red-black Gauss-Seidel smoothing for the multigrid preconditioner (an alternative to the
// reference's weighted Jacobi, multigrid.wgsl:66-73).  The cycle stays symmetric: the
// pre-smoother sweeps red then black from a zero guess, the post-smoother black then red.
// Red cells are those with x + y + z even; a level is eligible when each periodic axis has even
// length there (the colouring must survive the wrap).
//
// With unit weight every Gauss-Seidel update replaces the old value, so with u = r / d (the
// residual scaled by the stored reciprocal diagonal, 0 on inactive cells) the passes reduce to
//   pre-smoother   z_red = u,   z_black = u + dinv * sum_faces beta * z_red(neighbour)
//   residual       red: sum_faces beta * z_black(neighbour)  (black: 0, exactly)
//   post-smoother  v = u + s * zc(parent) on active red cells (u on black cells),
//                  t_black = u + dinv * sum beta * v(neighbour),
//                  t_red   = u + dinv * sum beta * t_black(neighbour)
// where "dinv" is the stored reciprocal diagonal: the smoother's diagonal is 1/dinv, used
// consistently in both sweeps and in the residual, so the sweeps stay each other's adjoint.
//
// One kernel shape serves both passes on any level: a block owns a tile of 64 x TY cells of every
// plane of a z-chunk and marches up z.  For each new plane p it stages, over the tile and a halo of
// two cells, the field u (or v) and the coefficients into LDS (stage 1); computes the black values
// of plane p-1 over the tile and a halo of one (stage 2); and finishes the red cells of plane p-2
// (stage 3): their restricted residual (summed over the 2x2x2 aggregates) or the smoothed output.
// The LDS rings hold three planes of the staged field, the coefficients and the black values; the
// loads of plane p + 1 are issued before stages 2 and 3 of plane p, so they are in flight while the
// block computes.  Arithmetic FP32; storage CT/VT (BF16 in the mixed solver, FP32 otherwise).
#include "mhdp_kernels.h"
#include "launch.h"

#include <type_traits>

#define CONVERGED(sc) ((sc)->conv > 0.5)

namespace {

constexpr int RB_TX = 64;
enum RbMode : int { RB_RESTRICT = 0, RB_SMOOTH = 1, RB_RESTRICT_UPDATE = 2 };

// Coordinate v of an axis of length n: itself in range, wrapped on a periodic axis, else -1.
__device__ __forceinline__ int rb_coord(int v, int n, bool per) {
    if (v >= 0 && v < n) return v;
    if (!per) return -1;
    v %= n;
    return v < 0 ? v + n : v;
}

template <int V>
using IC = std::integral_constant<int, V>;

// Coefficients travel through LDS as their raw storage bits (BF16: the upper half of an FP32).
template <class CT> struct RbRaw;
// ld(p, i): element i read through a 32-bit byte offset from the array's (uniform) base, so the
// load takes the scalar-base + 32-bit-offset form instead of 64-bit per-lane address arithmetic
// (every array here is far below 4 GiB).
template <> struct RbRaw<float> {
    using type = float;
    static __device__ __forceinline__ float ld(const float* p, uint32_t i) {
        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + (i << 2));
    }
    static __device__ __forceinline__ float f(float v) { return v; }
};
template <> struct RbRaw<__hip_bfloat16> {
    using type = unsigned short;
    static __device__ __forceinline__ unsigned short ld(const __hip_bfloat16* p, uint32_t i) {
        return *reinterpret_cast<const unsigned short*>(reinterpret_cast<const char*>(p) + (i << 1));
    }
    static __device__ __forceinline__ float f(unsigned short v) { return __uint_as_float((uint32_t)v << 16); }
};
template <class T>
__device__ __forceinline__ float rb_ldf(const T* p, uint32_t i) {
    return RbRaw<T>::f(RbRaw<T>::ld(p, i));
}
__device__ __forceinline__ float rb_ldq(const float* p, uint32_t i) { return RbRaw<float>::ld(p, i); }
template <class T>
__device__ __forceinline__ void rb_st(T* p, uint32_t i, float v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (uint32_t)sizeof(T)) = cvt<T>(v);
}

}  // namespace

// MODE RB_RESTRICT: pre-smoother and restriction of its residual into out (the coarse right-hand
// side).  RB_RESTRICT_UPDATE (fine level): the same after the PCG residual update
// r' = (r - mean) - alpha q (stored in the twin buffer, sums of r' and r'^2 -> FIN_RESTRICT_UPD),
// as the Jacobi passes do.
// MODE RB_SMOOTH: post-smoother after the coarse correction zc, output t (out), on the fine level
// the reduction of (r - mean) . t -> `mode`; on a coarse level with base set, out = base + theta t.
template <class CT, class VT, class RT, int MODE, bool COARSE, int TY>
struct K_rb {
    static constexpr int MINB = 2;
    static constexpr int RX = RB_TX + 4, RY = TY + 4;  // stage 1: tile + halo 2 (columns, rows)
    static constexpr int PX = RX / 2;                  // column pairs of a region row
    static constexpr int PN = PX * RY;                 // elements of one column-parity half-plane
    static constexpr int SB = (RB_TX + 2) / 2, SY = TY + 2, SN = SB * SY;  // stage 2: black positions
    static constexpr int NOWN = (RB_TX / 2) * (TY / 2);  // 2x2 footprints of the tile (one per thread)
    static constexpr int NH = 4 * RX + 4 * TY;           // halo positions of a plane
    static constexpr int HS = (NH + MHDP_BLOCK - 1) / MHDP_BLOCK;
    static constexpr int M2 = (SN + MHDP_BLOCK - 1) / MHDP_BLOCK;
    static_assert(NOWN <= MHDP_BLOCK, "one footprint per thread");
    struct Args {
        Grid3 g, cg;           // this level and the next coarser one
        const RT* r0;          // residual (fine: the buffer Scal.rsel names, r0 or r1)
        const RT* r1;
        float* w0;             // fine restriction with update: where r' goes (the other buffer)
        float* w1;
        const float* q;        // fine restriction with update: the PCG product
        const CT* dinv;
        const CT* bx;
        const CT* by;
        const CT* bz;
        const VT* zc;          // RB_SMOOTH: coarse correction
        VT* out;               // RB_RESTRICT: coarse right-hand side; RB_SMOOTH: t
        const VT* base;        // RB_SMOOTH coarse: out = base + theta * t when set
        float theta;
        double* part;
        Scal* sc;
        int mode;              // RB_SMOOTH fine: finalize of the (r - mean) . t reduction
        int use_next;          // fine: the cycle's mean is Scal.mean_next
        uint32_t zlen;         // planes per z-chunk (even)
    };
    using RR = RbRaw<CT>;
    using CS = typename RR::type;
    // LDS layout: every per-plane field is split by column parity into two half-planes of PN
    // elements, [parity][row * PX + column / 2].  Red and black cells alternate along a row, so a
    // stage that works on one colour reads its neighbours at unit stride (no bank conflicts).
    static __device__ __forceinline__ int lidx(int rx, int ry) { return (rx & 1) * PN + ry * PX + (rx >> 1); }

    static __device__ void run(const Args& A, const VGrid& vg) {
        Scal* sc = A.sc;
        if (CONVERGED(sc)) return;
        __shared__ float W[3][2 * PN + 1];
        __shared__ CS Cx[3][2 * PN + 1], Cy[3][2 * PN + 1], Cz[3][2 * PN + 1], Cd[2][2 * PN + 1];
        __shared__ float S2[3][SN + 1];  // black values of the stage-2 region, [row * SB + (rx - 1) / 2]
        constexpr int SPARE = 2 * PN;   // the slot unused positions write to
        const Grid3 g = A.g;
        const bool perx = g.per & 1u, pery = g.per & 2u, perz = g.per & 4u;
        constexpr bool upd = !COARSE && MODE == RB_RESTRICT_UPDATE;
        constexpr bool restrict_pass = MODE != RB_SMOOTH;
        // the mean the cycle's residual carries (deferred mean-free projection); after the update
        // the new residual's mean is Scal.mean_next
        const float m = COARSE ? 0.0f : (float)(A.use_next ? sc->mean_next : sc->mean);
        const float mn = COARSE ? 0.0f : (float)sc->mean_next;
        const float alpha = (float)sc->alpha;
        const float ps = (float)sc->pscale;
        const uint32_t rsel = COARSE ? 0u : sc->rsel;
        const RT* __restrict__ rin = rsel ? A.r1 : A.r0;
        float* __restrict__ rout = rsel ? A.w0 : A.w1;
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const uint32_t csy = (uint32_t)A.cg.nx, csz = (uint32_t)A.cg.nx * (uint32_t)A.cg.ny;
        const int txn = (g.nx + RB_TX - 1) / RB_TX, tyn = (g.ny + TY - 1) / TY;
        const uint32_t tiles = (uint32_t)txn * (uint32_t)tyn;
        const uint32_t items = tiles * (((uint32_t)g.nz + A.zlen - 1) / A.zlen);
        const int tid = (int)threadIdx.x;
        const bool has_own = tid < NOWN;
        const int ac = tid & (RB_TX / 2 - 1), ar = tid / (RB_TX / 2);  // this thread's 2x2 footprint
        double acc[4] = {0.0, 0.0, 0.0, 0.0};  // update: sum r'^2 in [1], sum r' in [3]
        double dot[1] = {0.0};                 // smoother: (r - mean) . t
        for (uint32_t it = vg.b; it < items; it += vg.n) {
            const uint32_t ch = it / tiles, rem = it - ch * tiles;
            const int ty = (int)(rem / (uint32_t)txn), tx = (int)(rem - (uint32_t)ty * (uint32_t)txn);
            const int x0 = tx * RB_TX, y0 = ty * TY;
            const int za = (int)(ch * A.zlen), zb = min(za + (int)A.zlen, g.nz);
            // ---- per-item tables
            // own footprint: cells (X, Y), (X + 1, Y), (X, Y + 1), (X + 1, Y + 1); region column 2 ac + 2
            // (in*: an output of this tile; src*: where the position's data comes from - past the grid
            // edge of a partial tile a periodic axis wraps, a closed one leaves the position absent)
            const int X = x0 + 2 * ac, Y = y0 + 2 * ar;
            const bool in00 = has_own && X < g.nx && Y < g.ny, in10 = in00 && X + 1 < g.nx;
            const bool in01 = has_own && X < g.nx && Y + 1 < g.ny, in11 = in01 && X + 1 < g.nx;
            const uint32_t o00 = in00 ? (uint32_t)X + sy * (uint32_t)Y : 0u;
            const uint32_t o01 = in01 ? (uint32_t)X + sy * (uint32_t)(Y + 1) : 0u;
            const uint32_t o10 = in10 ? o00 + 1u : 0u, o11 = in11 ? o01 + 1u : 0u;
            bool sok[4];
            uint32_t src[4], scpar[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int gx = rb_coord(X + (k & 1), g.nx, perx), gy = rb_coord(Y + (k >> 1), g.ny, pery);
                sok[k] = has_own && gx >= 0 && gy >= 0;
                src[k] = sok[k] ? (uint32_t)gx + sy * (uint32_t)gy : 0u;
                scpar[k] = sok[k] ? (uint32_t)(gx >> 1) + csy * (uint32_t)(gy >> 1) : 0u;  // parent (SMOOTH)
            }
            // halo slots: position (rx, ry) of the region, in-plane index (-1: absent), parent
            int hj[HS], hpos[HS], hpar[HS];
            uint32_t hcp[HS];
#pragma unroll
            for (int k = 0; k < HS; k++) {
                const int h = tid + MHDP_BLOCK * k;
                int rx = 0, ry = 0;
                hj[k] = SPARE;
                hpos[k] = -1;
                hcp[k] = 0;
                hpar[k] = 0;
                if (h < NH) {
                    if (h < 4 * RX) {
                        const int band = h / RX;  // rows 0, 1 and TY + 2, TY + 3
                        rx = h - band * RX;
                        ry = band < 2 ? band : TY + band;
                    } else {
                        const int h2 = h - 4 * RX, c = h2 & 3;
                        ry = 2 + (h2 >> 2);
                        rx = c < 2 ? c : RX - 4 + c;
                    }
                    hj[k] = lidx(rx, ry);
                    const int HX = x0 - 2 + rx, HY = y0 - 2 + ry;
                    hpar[k] = (HX + HY) & 1;
                    const int gx = rb_coord(HX, g.nx, perx), gy = rb_coord(HY, g.ny, pery);
                    if (gx >= 0 && gy >= 0) {
                        hpos[k] = gx + (int)sy * gy;
                        hcp[k] = (uint32_t)(gx >> 1) + csy * (uint32_t)(gy >> 1);
                    }
                }
            }
            // stage-2 positions: region row, and the half-plane index of the black cell on planes of
            // even / odd index (the black column of a row flips with the plane)
            int t2row[M2], t2e[M2], t2o[M2];
#pragma unroll
            for (int mm = 0; mm < M2; mm++) {
                const int k2 = min(tid + MHDP_BLOCK * mm, SN - 1);
                const int row = k2 / SB, cb = k2 - row * SB;
                const int ry = row + 1;
                // black at (rx, ry) on plane p1: (x0 - 2 + rx) + (y0 - 2 + ry) + p1 odd, rx = 2 cb + 1 or 2 cb + 2
                const int par = (y0 + ry) & 1;  // rx parity of the black cell on planes with p1 even is par ^ 1
                t2row[mm] = ry;
                t2e[mm] = par ? 2 * cb + 2 : 2 * cb + 1;  // p1 even: rx odd when (y0 + ry) even
                t2o[mm] = par ? 2 * cb + 1 : 2 * cb + 2;  // p1 odd
            }
            // own cells' (r - mean) and reciprocal diagonal of planes p - 1 (A) and p - 2 (B)
            float rtA[4], rtB[4], dA[4], dB[4];
#pragma unroll
            for (int k = 0; k < 4; k++) rtA[k] = rtB[k] = dA[k] = dB[k] = 0.0f;
            float sa = 0.0f;  // restriction: this footprint's sum over the current plane pair

            struct Own {
                float r[4], q[4];
                CS d[4], bx[4], by[4], bz[4];
                float zc[4];
            };
            struct Halo {
                float r, q;
                CS d, bx, by, bz;
                float zc;
            };
            Own own;
            Halo halo[HS];
            auto issue = [&](int p) {
                const int pz = rb_coord(p, g.nz, perz);
                const bool pin = pz >= 0;
                const uint32_t base = pin ? sz * (uint32_t)pz : 0u;
                const uint32_t cbase = pin ? csz * (uint32_t)(pz >> 1) : 0u;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t oi = pin && sok[k] ? src[k] + base : 0u;
                    own.r[k] = rb_ldf(rin, oi);
                    own.q[k] = upd ? rb_ldq(A.q, oi) : 0.0f;
                    own.d[k] = RR::ld(A.dinv, oi);
                    own.bx[k] = RR::ld(A.bx, oi);
                    own.by[k] = RR::ld(A.by, oi);
                    own.bz[k] = RR::ld(A.bz, oi);
                    // the four cells of an in-grid footprint share one parent (one cache line)
                    if (MODE == RB_SMOOTH) own.zc[k] = rb_ldf(A.zc, pin && sok[k] ? scpar[k] + cbase : 0u);
                }
#pragma unroll
                for (int k = 0; k < HS; k++) {
                    const bool ok = pin && hpos[k] >= 0;
                    const uint32_t i = ok ? (uint32_t)hpos[k] + base : 0u;
                    halo[k].r = rb_ldf(rin, i);
                    halo[k].q = upd ? rb_ldq(A.q, i) : 0.0f;
                    halo[k].d = RR::ld(A.dinv, i);
                    halo[k].bx = RR::ld(A.bx, i);
                    halo[k].by = RR::ld(A.by, i);
                    halo[k].bz = RR::ld(A.bz, i);
                    if (MODE == RB_SMOOTH) halo[k].zc = rb_ldf(A.zc, ok ? hcp[k] + cbase : 0u);
                }
            };
            __syncthreads();  // the previous item's last reads of the rings are done
            issue(za - 2);
            // One plane step; ring slots (p mod 3) and the plane's parity are compile-time constants.
            auto plane = [&](int p, auto S0c, auto CPc) {
                constexpr int s0 = decltype(S0c)::value, s1 = (s0 + 2) % 3, s2 = (s0 + 1) % 3;
                constexpr int cp = decltype(CPc)::value;  // p & 1
                // ---- stage 1: plane p over the tile and a halo of two, from the loads issued before
                float rtN[4], dN[4];
                {
                    const bool pin = rb_coord(p, g.nz, perz) >= 0;
                    const bool own_plane = p >= za && p < zb;
                    const uint32_t base = pin ? sz * (uint32_t)rb_coord(p, g.nz, perz) : 0u;
                    // the value staged at a cell: u = dinv (r - mean) (after the update on the fine
                    // level), plus the scaled coarse correction on active red cells (post-smoother)
                    auto stage = [&](bool ok, float r, float q, CS dr, float zc, bool red, float& rt_out,
                                     float& d_out, bool own_cell, uint32_t gi) {
                        const float d = ok ? RR::f(dr) : 0.0f;
                        float rt;
                        if constexpr (upd) {
                            const float rp = d != 0.0f ? (r - m) - alpha * q : 0.0f;
                            rt = rp - mn;
                            if (own_cell && ok && own_plane) {
                                rb_st(rout, gi, rp);
                                acc[1] += (double)rp * (double)rp;
                                acc[3] += (double)rp;
                            }
                        } else {
                            rt = r - m;
                        }
                        rt = ok ? rt : 0.0f;
                        rt_out = rt;
                        d_out = d;
                        float wv = d * rt;
                        if (MODE == RB_SMOOTH) wv += (red && d != 0.0f) ? ps * zc : 0.0f;
                        return wv;
                    };
                    if (has_own) {
                        const bool outs[4] = {in00, in10, in01, in11};
                        const uint32_t offs[4] = {o00, o10, o01, o11};
                        bool oks[4];
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int dx = k & 1, dy = k >> 1;
                            oks[k] = pin && sok[k];
                            const bool red = ((X + dx + Y + dy + p) & 1) == 0;
                            const float wv = stage(oks[k], own.r[k], own.q[k], own.d[k], own.zc[k], red, rtN[k],
                                                   dN[k], outs[k], offs[k] + base);
                            const int j = lidx(2 * ac + 2 + dx, 2 * ar + 2 + dy);
                            W[s0][j] = wv;
                            Cx[s0][j] = oks[k] ? own.bx[k] : CS(0);
                            Cy[s0][j] = oks[k] ? own.by[k] : CS(0);
                            Cz[s0][j] = oks[k] ? own.bz[k] : CS(0);
                            Cd[cp][j] = oks[k] ? own.d[k] : CS(0);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < HS; k++) {
                        const bool ok = pin && hpos[k] >= 0;
                        float rt_, d_;
                        const float wv = stage(ok, halo[k].r, halo[k].q, halo[k].d, halo[k].zc, ((hpar[k] + p) & 1) == 0,
                                               rt_, d_, false, 0u);
                        const int j = hj[k];
                        W[s0][j] = wv;
                        Cx[s0][j] = ok ? halo[k].bx : CS(0);
                        Cy[s0][j] = ok ? halo[k].by : CS(0);
                        Cz[s0][j] = ok ? halo[k].bz : CS(0);
                        Cd[cp][j] = ok ? halo[k].d : CS(0);
                    }
                }
                if (p + 1 <= zb + 1) issue(p + 1);  // in flight during stages 2 and 3
                __syncthreads();
                // ---- stage 2: black values of plane p1 = p - 1 over the tile and a halo of one
                const int p1 = p - 1;
                if (p1 >= za - 1 && p1 <= zb) {
#pragma unroll
                    for (int mm = 0; mm < M2; mm++) {
                        const int k2 = tid + MHDP_BLOCK * mm;
                        if (M2 * MHDP_BLOCK > SN && k2 >= SN) break;  // the last round: the first waves only
                        const int ry = t2row[mm];
                        const int rx = cp ? t2e[mm] : t2o[mm];  // cp = p & 1: p1 even when p odd
                        const int c = lidx(rx, ry);
                        const int xl = lidx(rx - 1, ry), xu = lidx(rx + 1, ry);
                        const float d = RR::f(Cd[cp ^ 1][c]);
                        float sum = 0.0f;
                        sum += RR::f(Cx[s1][c]) * W[s1][xl];
                        sum += RR::f(Cx[s1][xu]) * W[s1][xu];
                        sum += RR::f(Cy[s1][c]) * W[s1][c - PX];
                        sum += RR::f(Cy[s1][c + PX]) * W[s1][c + PX];
                        sum += RR::f(Cz[s1][c]) * W[s2][c];
                        sum += RR::f(Cz[s0][c]) * W[s0][c];
                        S2[s1][k2] = W[s1][c] + d * sum;
                    }
                }
                __syncthreads();
                // ---- stage 3: this thread's footprint on plane q = p - 2; black values: q in s2,
                // q + 1 in s1, q - 1 in s0
                const int q = p - 2;
                if (has_own && q >= za && q < zb) {
                    float t[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int dx = k & 1, dy = k >> 1;
                        const int rx = 2 * ac + 2 + dx, ry = 2 * ar + 2 + dy;
                        const bool red = ((X + dx + Y + dy + q) & 1) == 0;
                        // black-value index of stage-2 position (rx', ry'): (ry' - 1) * SB + (rx' - 1) / 2
                        const int sc_ = (ry - 1) * SB + ((rx - 1) >> 1);
                        if (red) {
                            const int c = lidx(rx, ry);
                            const float d = dB[k];
                            float sum = 0.0f;
                            sum += RR::f(Cx[s2][c]) * S2[s2][(ry - 1) * SB + ((rx - 2) >> 1)];
                            sum += RR::f(Cx[s2][lidx(rx + 1, ry)]) * S2[s2][(ry - 1) * SB + (rx >> 1)];
                            sum += RR::f(Cy[s2][c]) * S2[s2][sc_ - SB];
                            sum += RR::f(Cy[s2][c + PX]) * S2[s2][sc_ + SB];
                            sum += RR::f(Cz[s2][c]) * S2[s0][sc_];
                            sum += RR::f(Cz[s1][c]) * S2[s1][sc_];
                            t[k] = restrict_pass ? (d != 0.0f ? sum : 0.0f) : d * rtB[k] + d * sum;
                        } else {
                            t[k] = restrict_pass ? 0.0f : S2[s2][sc_];
                        }
                    }
                    const bool ins[4] = {in00, in10, in01, in11};
                    if constexpr (restrict_pass) {
#pragma unroll
                        for (int k = 0; k < 4; k++) sa += ins[k] ? t[k] : 0.0f;
                        if ((q & 1) == 1 || q == g.nz - 1) {  // both planes of the aggregate layer are in
                            if (in00)
                                rb_st(A.out, (uint32_t)(X >> 1) + csy * (uint32_t)(Y >> 1) + csz * (uint32_t)(q >> 1), sa);
                            sa = 0.0f;
                        }
                    } else {
                        const uint32_t qb = sz * (uint32_t)q;
                        const uint32_t offs[4] = {o00, o10, o01, o11};
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            if (!ins[k]) continue;
                            const uint32_t i = offs[k] + qb;
                            float v = t[k];
                            if (COARSE && A.base) v = rb_ldf(A.base, i) + A.theta * v;
                            rb_st(A.out, i, v);
                            if (!COARSE && dB[k] != 0.0f) dot[0] += (double)rtB[k] * (double)v;
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    rtB[k] = rtA[k];
                    rtA[k] = rtN[k];
                    dB[k] = dA[k];
                    dA[k] = dN[k];
                }
                __syncthreads();  // stage 3's reads of plane q's coefficients precede their overwrite
            };
            for (int p = za - 2; p <= zb + 1; p++) {
                switch (((p % 6) + 6) % 6) {
                case 0: plane(p, IC<0>(), IC<0>()); break;
                case 1: plane(p, IC<1>(), IC<1>()); break;
                case 2: plane(p, IC<2>(), IC<0>()); break;
                case 3: plane(p, IC<0>(), IC<1>()); break;
                case 4: plane(p, IC<1>(), IC<0>()); break;
                default: plane(p, IC<2>(), IC<1>()); break;
                }
            }
        }
        if constexpr (!COARSE) {
            if constexpr (upd) reduce_finish<4>(vg, acc, A.part, sc, 3, FIN_RESTRICT_UPD);
            else if constexpr (MODE == RB_SMOOTH) reduce_finish<1>(vg, dot, A.part, sc, 4, A.mode);
        }
    }
};

// ---------------------------------------------------------------- launchers

bool rb_level_ok(const Grid3& g) {
    for (int a = 0; a < 3; a++) {
        const int n = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
        if (((g.per >> a) & 1u) && (n & 1)) return false;
    }
    return true;
}

// Planes per z-chunk: long enough that the four planes of halo stay a small part of a chunk, short
// enough that the launch has some thousand items (tiles x chunks); even, 8..64.
static uint32_t rb_zlen(const Grid3& g, int ty) {
    const uint64_t tiles = (uint64_t)((g.nx + RB_TX - 1) / RB_TX) * (uint64_t)((g.ny + ty - 1) / ty);
    uint64_t z = (uint64_t)g.nz * tiles / 1024u;
    z &= ~(uint64_t)1;
    return (uint32_t)(z < 8 ? 8 : (z > 64 ? 64 : z));
}
template <class K>
static uint32_t rb_blocks(const Grid3& g, uint32_t zlen, int ty) {
    const uint32_t tiles = (uint32_t)((g.nx + RB_TX - 1) / RB_TX) * (uint32_t)((g.ny + ty - 1) / ty);
    const uint32_t items = tiles * (((uint32_t)g.nz + zlen - 1) / zlen);
    return grid_blocks(items * MHDP_BLOCK, launch_tuning().march_cap);
}

template <class CT, class VT, class RT, int MODE, bool COARSE, int TY>
static void rb_go(hipStream_t s, const typename K_rb<CT, VT, RT, MODE, COARSE, TY>::Args& a0) {
    using K = K_rb<CT, VT, RT, MODE, COARSE, TY>;
    typename K::Args a = a0;
    a.zlen = rb_zlen(a.g, TY);
    go<K>(s, rb_blocks<K>(a.g, a.zlen, TY), a);
}

void launch_rb_restrict(hipStream_t s, int store, bool coarse, const Grid3& cg, const Grid3& g, const void* r,
                        const void* r2, float* w0, float* w1, const float* q, const void* dinv, const void* bx,
                        const void* by, const void* bz, void* rc, double* part, Scal* sc) {
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        if (coarse)
            rb_go<T, T, T, RB_RESTRICT, true, 16>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                    (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, nullptr,
                                                    (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else if (q)
            rb_go<T, T, float, RB_RESTRICT_UPDATE, false, 16>(
                s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q, (const T*)dinv, (const T*)bx, (const T*)by,
                    (const T*)bz, nullptr, (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_RESTRICT, false, 16>(s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q,
                                                         (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz,
                                                         nullptr, (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
    } else {
        using T = float;
        if (coarse)
            rb_go<T, T, T, RB_RESTRICT, true, 8>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                   (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, nullptr,
                                                   (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else if (q)
            rb_go<T, T, float, RB_RESTRICT_UPDATE, false, 8>(
                s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q, (const T*)dinv, (const T*)bx, (const T*)by,
                    (const T*)bz, nullptr, (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_RESTRICT, false, 8>(s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q,
                                                        (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz,
                                                        nullptr, (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
    }
}

void launch_rb_smooth(hipStream_t s, int store, bool coarse, const Grid3& g, const Grid3& cg, const void* r,
                      const void* r2, const void* dinv, const void* bx, const void* by, const void* bz, const void* zc,
                      void* t, const void* base, float theta, double* part, Scal* sc, int mode, bool use_next) {
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        if (coarse)
            rb_go<T, T, T, RB_SMOOTH, true, 16>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                  (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz,
                                                  (const T*)zc, (T*)t, (const T*)base, theta, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_SMOOTH, false, 16>(s, {g, cg, (const float*)r, (const float*)r2, nullptr, nullptr,
                                                       nullptr, (const T*)dinv, (const T*)bx, (const T*)by,
                                                       (const T*)bz, (const T*)zc, (T*)t, nullptr, 0.0f, part, sc,
                                                       mode, use_next ? 1 : 0, 0});
    } else {
        using T = float;
        if (coarse)
            rb_go<T, T, T, RB_SMOOTH, true, 8>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                 (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz,
                                                 (const T*)zc, (T*)t, (const T*)base, theta, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_SMOOTH, false, 8>(s, {g, cg, (const float*)r, (const float*)r2, nullptr, nullptr,
                                                      nullptr, (const T*)dinv, (const T*)bx, (const T*)by,
                                                      (const T*)bz, (const T*)zc, (T*)t, nullptr, 0.0f, part, sc,
                                                      mode, use_next ? 1 : 0, 0});
    }
}
