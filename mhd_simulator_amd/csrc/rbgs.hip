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
// One kernel shape serves both passes on any level (K_rb below): a wave owns a strip of 64
// columns by a few rows of every plane of a z-chunk and marches up z, with the rows in registers,
// x neighbours from DPP lane shifts and no LDS.  Arithmetic FP32; storage CT/VT (BF16 in the
// mixed solver, FP32 otherwise).
#include "mhdp_kernels.h"
#include "launch.h"
#include "mob.h"

#include <type_traits>

#define CONVERGED(sc) ((sc)->conv > 0.5)

namespace {

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

// Whole-wave DPP lane moves (full EXEC only): the value of lane - 1, lane + 1, lane ^ 1 (0 where
// the source lane does not exist).
__device__ __forceinline__ float rb_prev_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float rb_next_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float rb_pair_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
}
// Raw buffer access: a descriptor per array (base, size in bytes), a per-lane byte offset and a
// uniform (scalar) byte offset.  A lane offset past the array's size reads 0 (the buffer unit's
// range check), which is how absent columns read zeros.
using RbRsrc = __amdgpu_buffer_rsrc_t;
constexpr uint32_t RB_ABSENT = 0x7FFFFFF0u;  // lane offset past every array (all are below 2 GiB)
__device__ __forceinline__ RbRsrc rb_rsrc(const void* p, uint64_t bytes) {
    // word 3 = 0x00020000: the gfx9 untyped-buffer format
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <class T> struct RbBuf;
template <> struct RbBuf<float> {
    using raw = float;
    static constexpr uint32_t size = 4;
    static __device__ __forceinline__ float ld(RbRsrc R, uint32_t v, uint32_t s) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(R, (int)v, (int)s, 0));
    }
    static __device__ __forceinline__ float f(float v) { return v; }
    static __device__ __forceinline__ void st(RbRsrc R, float x, uint32_t v, uint32_t s) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), R, (int)v, (int)s, 0);
    }
};
template <> struct RbBuf<__hip_bfloat16> {
    using raw = unsigned short;
    static constexpr uint32_t size = 2;
    static __device__ __forceinline__ unsigned short ld(RbRsrc R, uint32_t v, uint32_t s) {
        return __builtin_amdgcn_raw_buffer_load_b16(R, (int)v, (int)s, 0);
    }
    static __device__ __forceinline__ float f(unsigned short v) { return __uint_as_float((uint32_t)v << 16); }
    static __device__ __forceinline__ void st(RbRsrc R, float x, uint32_t v, uint32_t s) {
        const __hip_bfloat16 b = __float2bfloat16(x);
        unsigned short bits;
        __builtin_memcpy(&bits, &b, 2);
        __builtin_amdgcn_raw_buffer_store_b16(bits, R, (int)v, (int)s, 0);
    }
};
template <class T>
__device__ __forceinline__ float rb_bf(RbRsrc R, uint32_t v, uint32_t s) {
    return RbBuf<T>::f(RbBuf<T>::ld(R, v, s));
}

constexpr int RBW_OUT = 60;  // output columns of a 64-lane strip (two halo lanes each side)

}  // namespace

// MODE RB_RESTRICT: pre-smoother and restriction of its residual into out (the coarse right-hand
// side).  RB_RESTRICT_UPDATE (fine level): the same after the PCG residual update
// r' = (r - mean) - alpha q (stored in the twin buffer, sums of r' and r'^2 -> FIN_RESTRICT_UPD),
// as the Jacobi passes do.
// MODE RB_SMOOTH: post-smoother after the coarse correction zc, output t (out), on the fine level
// the reduction of (r - mean) . t -> `mode`; on a coarse level with base set, out = base + theta t.
//
// One wave per work item and no LDS: a wave owns a strip of 64 columns (lane = column; the middle
// 60 are its outputs, two halo lanes each side) by H rows of every plane of a z-chunk, and marches
// up z.  Rows are registers, x neighbours come from DPP lane shifts, z
// neighbours from a three-plane register ring.  Step p stages plane p over the strip's rows and a
// halo of two rows (stage 1: u or v), computes plane p - 1's black values over the rows and a halo
// of one (stage 2) and finishes plane p - 2's red cells (stage 3).  The waves of a block take
// separate items and never wait for each other.  Both colours are computed on every lane and the
// colour selects: one instruction stream for the wave.
template <class CT, class VT, class RT, int MODE, bool COARSE, int H>
struct K_rb {
    static constexpr int MINB = 3;  // waves per SIMD the registers allow
    static constexpr int NR = H + 4;  // stage-1 rows -2 .. H + 1
    static constexpr int N2 = H + 2;  // stage-2 rows -1 .. H
    static_assert(H % 2 == 0, "row pairs of the restriction");
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

    static __device__ void run(const Args& A, const VGrid& vg) {
        Scal* sc = A.sc;
        if (CONVERGED(sc)) return;
        const Grid3 g = A.g;
        const bool perx = g.per & 1u, pery = g.per & 2u, perz = g.per & 4u;
        constexpr bool upd = !COARSE && MODE == RB_RESTRICT_UPDATE;
        constexpr bool restrict_pass = MODE != RB_SMOOTH;
        const float m = COARSE ? 0.0f : (float)(A.use_next ? sc->mean_next : sc->mean);
        const float mn = COARSE ? 0.0f : (float)sc->mean_next;
        const float alpha = (float)sc->alpha;
        const float ps = (float)sc->pscale;
        const uint32_t rsel = COARSE ? 0u : sc->rsel;
        const RT* __restrict__ rin = rsel ? A.r1 : A.r0;
        float* __restrict__ rout = rsel ? A.w0 : A.w1;
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const uint32_t csy = (uint32_t)A.cg.nx, csz = (uint32_t)A.cg.nx * (uint32_t)A.cg.ny;
        const uint32_t nsx = ((uint32_t)g.nx + RBW_OUT - 1) / RBW_OUT, nty = ((uint32_t)g.ny + H - 1) / H;
        const uint32_t tiles = nsx * nty;
        const uint32_t items = tiles * (((uint32_t)g.nz + A.zlen - 1) / A.zlen);
        const uint64_t cells = (uint64_t)sz * (uint64_t)g.nz;
        const uint64_t ccells = (uint64_t)csz * (uint64_t)A.cg.nz;
        const RbRsrc Rr = rb_rsrc(rin, cells * RbBuf<RT>::size);
        const RbRsrc Rq = rb_rsrc(upd ? (const void*)A.q : (const void*)rin, cells * 4u);
        const RbRsrc Rd = rb_rsrc(A.dinv, cells * RbBuf<CT>::size);
        const RbRsrc Rx = rb_rsrc(A.bx, cells * RbBuf<CT>::size);
        const RbRsrc Ry = rb_rsrc(A.by, cells * RbBuf<CT>::size);
        const RbRsrc Rz = rb_rsrc(A.bz, cells * RbBuf<CT>::size);
        const RbRsrc Rc = rb_rsrc(MODE == RB_SMOOTH ? (const void*)A.zc : (const void*)A.dinv,
                                  (MODE == RB_SMOOTH ? ccells : cells) * RbBuf<VT>::size);
        const RbRsrc Rw = rb_rsrc(upd ? (const void*)rout : (const void*)A.out,
                                  upd ? cells * 4u : (restrict_pass ? ccells : cells) * RbBuf<VT>::size);
        const RbRsrc Ro = rb_rsrc(A.out, (restrict_pass ? ccells : cells) * RbBuf<VT>::size);
        const RbRsrc Rb = rb_rsrc(COARSE && A.base ? (const void*)A.base : (const void*)A.out, cells * RbBuf<VT>::size);
        const int lane = (int)(threadIdx.x & 63u);
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint32_t waves = (uint32_t)MHDP_BLOCK / 64u;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};  // update: sum r'^2 in [1], sum r' in [3]
        double dot[1] = {0.0};                 // smoother: (r - mean) . t
        for (uint32_t it = vg.b * waves + wave; it < items; it += vg.n * waves) {
            const uint32_t ch = it / tiles, rem = it - ch * tiles;
            const uint32_t ty = rem / nsx, sx = rem - ty * nsx;
            const int x = (int)sx * RBW_OUT - 2 + lane, y0 = (int)ty * H;
            const int za = (int)(ch * A.zlen), zb = min(za + (int)A.zlen, g.nz);
            const int gx = rb_coord(x, g.nx, perx);
            const bool lane_ok = gx >= 0;
            // lane byte offsets of the load column (past the arrays when the column is absent)
            const uint32_t vR = lane_ok ? (uint32_t)gx * RbBuf<RT>::size : RB_ABSENT;
            const uint32_t vQ = lane_ok ? (uint32_t)gx * 4u : RB_ABSENT;
            const uint32_t vC = lane_ok ? (uint32_t)gx * RbBuf<CT>::size : RB_ABSENT;
            const uint32_t vZ = lane_ok ? (uint32_t)(gx >> 1) * RbBuf<VT>::size : RB_ABSENT;  // parent (SMOOTH)
            const bool own_lane = lane >= 2 && lane < 2 + RBW_OUT && x < g.nx;
            const bool lodd = ((x + y0) & 1) != 0;                     // parity of (x, y0) on plane 0
            // rows y0 - 2 + j: in-plane row offset (a safe row if absent), presence, parent row offset
            uint32_t roff[NR], rcoff[NR];
            bool rok[NR];
#pragma unroll
            for (int j = 0; j < NR; j++) {
                const int gy = rb_coord(y0 - 2 + j, g.ny, pery);
                rok[j] = gy >= 0;
                roff[j] = rok[j] ? sy * (uint32_t)gy : 0u;
                rcoff[j] = rok[j] ? csy * (uint32_t)(gy >> 1) : 0u;
            }
            // rings (slot = plane mod 3): staged field, black values, coefficients (stage-2 rows;
            // by one row more: the upper face of the last), own rows' (r - mean)
            float U[3][NR], W[3][N2];
            float CX[3][N2], CY[3][N2 + 1], CZ[3][N2], CD[3][N2];
            float RTO[3][H];
            float sa[H / 2];
#pragma unroll
            for (int i = 0; i < H / 2; i++) sa[i] = 0.0f;
            // ring slots count planes from the chunk's first staged plane za - 2
            auto step = [&](int p, auto S0c) {
                constexpr int s0 = decltype(S0c)::value, s1 = (s0 + 2) % 3, s2 = (s0 + 1) % 3;
                // colour masks of plane p: ev = (x + y0 + p even), the parity of row offset 0
                const bool ev = lodd == ((p & 1) != 0);
                // ---- stage 1: plane p, rows -2 .. H + 1
                {
                    const int pz = rb_coord(p, g.nz, perz);
                    const bool pok = pz >= 0;
                    const uint32_t poff = pok ? sz * (uint32_t)pz : 0u;
                    const uint32_t pcoff = pok ? csz * (uint32_t)(pz >> 1) : 0u;
                    const bool own_plane = p >= za && p < zb;
                    float rr[NR], qq[NR], dd[NR], zc[NR];
                    typename RbBuf<CT>::raw bx[N2], by[N2 + 1], bz[N2];
#pragma unroll
                    for (int j = 0; j < NR; j++) {
                        const uint32_t ro = roff[j] + poff;
                        rr[j] = rb_bf<RT>(Rr, vR, ro * RbBuf<RT>::size);
                        qq[j] = upd ? RbBuf<float>::ld(Rq, vQ, ro * 4u) : 0.0f;
                        dd[j] = rb_bf<CT>(Rd, vC, ro * RbBuf<CT>::size);
                        if (j >= 1 && j <= N2) {
                            bx[j - 1] = RbBuf<CT>::ld(Rx, vC, ro * RbBuf<CT>::size);
                            bz[j - 1] = RbBuf<CT>::ld(Rz, vC, ro * RbBuf<CT>::size);
                        }
                        if (j >= 1) by[j - 1] = RbBuf<CT>::ld(Ry, vC, ro * RbBuf<CT>::size);
                        if (MODE == RB_SMOOTH) zc[j] = rb_bf<VT>(Rc, vZ, (rcoff[j] + pcoff) * RbBuf<VT>::size);
                    }
#pragma unroll
                    for (int j = 0; j < NR; j++) {
                        const bool ok = rok[j] && pok;  // uniform: an absent column read zeros
                        const float d = ok ? dd[j] : 0.0f;
                        float rt;
                        if constexpr (upd) {
                            const float rp = d != 0.0f ? (rr[j] - m) - alpha * qq[j] : 0.0f;
                            rt = rp - mn;
                            if (j >= 2 && j < 2 + H && own_plane && own_lane && y0 + j - 2 < g.ny) {
                                RbBuf<float>::st(Rw, rp, (uint32_t)x * 4u, (sy * (uint32_t)(y0 + j - 2) + sz * (uint32_t)p) * 4u);
                                acc[1] += (double)rp * (double)rp;
                                acc[3] += (double)rp;
                            }
                        } else {
                            rt = rr[j] - m;
                        }
                        rt = ok ? rt : 0.0f;
                        float u = d * rt;
                        if constexpr (MODE == RB_SMOOTH) {
                            const bool red = (j & 1) ? !ev : ev;  // x + y0 - 2 + j + p even
                            u += (red && d != 0.0f) ? ps * zc[j] : 0.0f;
                            if (j >= 2 && j < 2 + H) RTO[s0][j - 2] = rt;
                        }
                        U[s0][j] = u;
                        // (an absent row's or plane's own values meet d = 0; its faces toward a
                        // present neighbour, the y and z lower slots, must read 0)
                        if (j >= 1 && j <= N2) {
                            CD[s0][j - 1] = d;
                            CX[s0][j - 1] = RbBuf<CT>::f(bx[j - 1]);
                            CZ[s0][j - 1] = ok ? RbBuf<CT>::f(bz[j - 1]) : 0.0f;
                        }
                        if (j >= 1) CY[s0][j - 1] = ok ? RbBuf<CT>::f(by[j - 1]) : 0.0f;
                    }
                }
                // ---- stage 2: black values of plane c = p - 1, rows -1 .. H (index k)
                if (p >= za) {
#pragma unroll
                    for (int k = 0; k < N2; k++) {
                        const float uc = U[s1][k + 1];
                        float s = CX[s1][k] * rb_prev_lane(uc);
                        s += rb_next_lane(CX[s1][k]) * rb_next_lane(uc);
                        s += CY[s1][k] * U[s1][k];
                        s += CY[s1][k + 1] * U[s1][k + 2];
                        s += CZ[s1][k] * U[s2][k + 1];
                        s += CZ[s0][k] * U[s0][k + 1];
                        // (x, y0 - 1 + k, p - 1) black: x + y0 + k + p odd
                        const bool black = (k & 1) ? ev : !ev;
                        W[s1][k] = black ? uc + CD[s1][k] * s : uc;
                    }
                }
                // ---- stage 3: red cells of plane e = p - 2, rows 0 .. H - 1 (i); black values of
                // e - 1 in slot s0, e in s2, e + 1 in s1
                const int e = p - 2;
                if (e >= za) {
                    float t[H];
#pragma unroll
                    for (int i = 0; i < H; i++) {
                        const int k = i + 1;
                        const float wc = W[s2][k];
                        float s = CX[s2][k] * rb_prev_lane(wc);
                        s += rb_next_lane(CX[s2][k]) * rb_next_lane(wc);
                        s += CY[s2][k] * W[s2][k - 1];
                        s += CY[s2][k + 1] * W[s2][k + 1];
                        s += CZ[s2][k] * W[s0][k];
                        s += CZ[s1][k] * W[s1][k];
                        // (x, y0 + i, e) red: x + y0 + i + p even
                        const bool red = (i & 1) ? !ev : ev;
                        const float d = CD[s2][k];
                        if constexpr (restrict_pass)
                            t[i] = (red && d != 0.0f) ? s : 0.0f;
                        else
                            t[i] = red ? d * RTO[s2][i] + d * s : wc;
                    }
                    if constexpr (restrict_pass) {
                        // aggregate sums: lane pairs (even x first), row pairs, plane pairs
#pragma unroll
                        for (int i = 0; i < H; i++) t[i] = (own_lane && y0 + i < g.ny) ? t[i] : 0.0f;
#pragma unroll
                        for (int i = 0; i < H / 2; i++) {
                            const float v = t[2 * i] + t[2 * i + 1];
                            sa[i] += v + rb_pair_lane(v);
                        }
                        if ((e & 1) == 1 || e == g.nz - 1) {
                            if (own_lane && (x & 1) == 0) {
#pragma unroll
                                for (int i = 0; i < H / 2; i++)
                                    if (y0 + 2 * i < g.ny)
                                        RbBuf<VT>::st(Ro, sa[i], (uint32_t)(x >> 1) * RbBuf<VT>::size,
                                                      (csy * (uint32_t)((y0 >> 1) + i) + csz * (uint32_t)(e >> 1)) *
                                                          RbBuf<VT>::size);
                            }
#pragma unroll
                            for (int i = 0; i < H / 2; i++) sa[i] = 0.0f;
                        }
                    } else if (own_lane) {
#pragma unroll
                        for (int i = 0; i < H; i++) {
                            if (y0 + i >= g.ny) continue;
                            const uint32_t vo = (uint32_t)x * RbBuf<VT>::size;
                            const uint32_t so = (sy * (uint32_t)(y0 + i) + sz * (uint32_t)e) * RbBuf<VT>::size;
                            float v = t[i];
                            if (COARSE && A.base) v = rb_bf<VT>(Rb, vo, so) + A.theta * v;
                            RbBuf<VT>::st(Ro, v, vo, so);
                            if (!COARSE && CD[s2][i + 1] != 0.0f) dot[0] += (double)RTO[s2][i] * (double)v;
                        }
                    }
                }
            };
            for (int p = za - 2; p <= zb + 1; p += 3) {
                step(p, IC<0>());
                if (p + 1 <= zb + 1) step(p + 1, IC<1>());
                if (p + 2 <= zb + 1) step(p + 2, IC<2>());
            }
        }
        if constexpr (!COARSE) {
            if constexpr (upd) reduce_finish<4>(vg, acc, A.part, sc, 3, FIN_RESTRICT_UPD);
            else if constexpr (MODE == RB_SMOOTH) reduce_finish<1>(vg, dot, A.part, sc, 4, A.mode);
        }
    }
};

// ---------------------------------------------------------------- pair-column kernel
//
// The same passes with the work laid out for fewer, wider loads and no halo rows in registers.
// A lane owns a PAIR of columns (x even, x + 1): on every row and plane one of the two is red and
// the other black, so each stage computes exactly one cell per lane, and the restriction's x pair
// is the lane itself.  A block of 1024 threads (16 waves) owns a tile of 32 pairs by 32 rows
// (64 x 32 columns, of which the middle 60 x 28 are outputs) of every plane of a z-chunk and
// marches up z; each wave holds two rows (half-waves of 32 lanes).  x neighbours come from DPP
// lane shifts, y neighbours from the LDS copy of the plane's values, z neighbours from the
// lane's own register ring.  Per plane step p: stage 1 stages plane p (u, coefficient pairs) and
// publishes the red values and the y-face pairs; stage 2 computes plane p - 1's black values and
// publishes them; stage 3 finishes plane p - 2's red cells.  Double-buffered LDS planes make one
// barrier per step enough.  Needs nx even (a pair never straddles a row end).
typedef unsigned int rb_v2u __attribute__((ext_vector_type(2)));
template <class T> struct RbP;
template <> struct RbP<float> {
    using raw = rb_v2u;
    static constexpr uint32_t esize = 4;
    static __device__ __forceinline__ raw ld(RbRsrc R, uint32_t v, uint32_t s) {
        return __builtin_amdgcn_raw_buffer_load_b64(R, (int)v, (int)s, 0);
    }
    static __device__ __forceinline__ float lo(raw r) { return __uint_as_float(r.x); }
    static __device__ __forceinline__ float hi(raw r) { return __uint_as_float(r.y); }
    static __device__ __forceinline__ raw zero() { return raw{0u, 0u}; }
    static __device__ __forceinline__ void st(RbRsrc R, float a, float b, uint32_t v, uint32_t s) {
        raw o;
        o.x = __float_as_uint(a);
        o.y = __float_as_uint(b);
        __builtin_amdgcn_raw_buffer_store_b64(o, R, (int)v, (int)s, 0);
    }
};
template <> struct RbP<__hip_bfloat16> {
    using raw = uint32_t;
    static constexpr uint32_t esize = 2;
    static __device__ __forceinline__ raw ld(RbRsrc R, uint32_t v, uint32_t s) {
        return __builtin_amdgcn_raw_buffer_load_b32(R, (int)v, (int)s, 0);
    }
    static __device__ __forceinline__ float lo(raw r) { return __uint_as_float(r << 16); }
    static __device__ __forceinline__ float hi(raw r) { return __uint_as_float(r & 0xFFFF0000u); }
    static __device__ __forceinline__ raw zero() { return 0u; }
    static __device__ __forceinline__ void st(RbRsrc R, float a, float b, uint32_t v, uint32_t s) {
        const __hip_bfloat16 ba = __float2bfloat16(a), bb = __float2bfloat16(b);
        unsigned short ua, ub;
        __builtin_memcpy(&ua, &ba, 2);
        __builtin_memcpy(&ub, &bb, 2);
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)ua | ((uint32_t)ub << 16), R, (int)v, (int)s, 0);
    }
};
// The coarse correction at a lane's parent (RB_SMOOTH), by element index.  BF16: the aligned dword
// holding it (a 16-bit load in the plane loop makes the compiler wait for it at once and copy the
// load ring at the loop's back edge); the array's range is rounded up to whole dwords.
template <class T> struct RbPar;
template <> struct RbPar<float> {
    using raw = float;
    static __device__ __forceinline__ raw ld(RbRsrc R, uint32_t e) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(R, (int)(e * 4u), 0, 0));
    }
    static __device__ __forceinline__ float f(raw r, uint32_t) { return r; }
};
template <> struct RbPar<__hip_bfloat16> {
    using raw = uint32_t;
    static __device__ __forceinline__ raw ld(RbRsrc R, uint32_t e) {
        return __builtin_amdgcn_raw_buffer_load_b32(R, (int)((e & ~1u) * 2u), 0, 0);
    }
    static __device__ __forceinline__ float f(raw r, uint32_t e) {
        return __uint_as_float((e & 1u) ? (r & 0xFFFF0000u) : (r << 16));
    }
};
constexpr uint32_t RBPAR_ABSENT = 0x1FFFFFF0u;  // parent index past every coarse array

template <class T>
__device__ __forceinline__ float rbp_sel(typename RbP<T>::raw r, bool hi) {
    return hi ? RbP<T>::hi(r) : RbP<T>::lo(r);
}

constexpr int RBP_LX = 32;                 // pairs per tile row (a half wave)
constexpr int RBP_OUTX = 2 * (RBP_LX - 2);  // 60 output columns

// NY: rows per tile (two per wave; NY - 4 output rows): 32 (1024 threads, one block per CU) for
// large levels, 16 (512 threads, two blocks per CU) for small ones, where a block's serial
// z-march is the latency to hide.
template <class CT, class VT, class RT, int MODE, bool COARSE, int NY>
struct K_rbp {
    static constexpr int RBP_NY = NY;
    static constexpr int RBP_OUTY = NY - 4;
    static constexpr int RBP_THREADS = RBP_LX * NY;
    static constexpr int THREADS = RBP_THREADS;
    static constexpr int MINB = NY == 32 ? 1 : 2;
    using Args = typename K_rb<CT, VT, RT, MODE, COARSE, 4>::Args;
    using PC = RbP<CT>;
    using PR = RbP<RT>;
    using PV = RbP<VT>;
    // one plane's loads of a lane (kept raw: a conversion placed early by the compiler would wait
    // for the load)
    struct Raw {
        typename PR::raw r;
        rb_v2u q;
        typename PC::raw d, bx, by, bz;
        typename RbPar<VT>::raw zc;
        typename PV::raw base;
    };

    static __device__ void run(const Args& A, const VGrid& vg) {
        Scal* sc = A.sc;
        if (CONVERGED(sc)) return;
        constexpr bool upd = !COARSE && MODE == RB_RESTRICT_UPDATE;
        constexpr bool restrict_pass = MODE != RB_SMOOTH;
        __shared__ float sU[2][RBP_NY][RBP_LX];                // red values of a plane (stage 1)
        __shared__ float sW[2][RBP_NY][RBP_LX];                // black values of a plane (stage 2)
        __shared__ typename PC::raw sB[3][RBP_NY][RBP_LX];     // y-face pairs of a plane
        const Grid3 g = A.g;
        const bool perx = g.per & 1u, pery = g.per & 2u, perz = g.per & 4u;
        const float m = COARSE ? 0.0f : (float)(A.use_next ? sc->mean_next : sc->mean);
        const float mn = COARSE ? 0.0f : (float)sc->mean_next;
        const float alpha = (float)sc->alpha;
        const float ps = (float)sc->pscale;
        const uint32_t rsel = COARSE ? 0u : sc->rsel;
        const RT* __restrict__ rin = rsel ? A.r1 : A.r0;
        float* __restrict__ rout = rsel ? A.w0 : A.w1;
        const uint32_t sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const uint32_t csy = (uint32_t)A.cg.nx, csz = (uint32_t)A.cg.nx * (uint32_t)A.cg.ny;
        const uint32_t ntx = ((uint32_t)g.nx + RBP_OUTX - 1) / RBP_OUTX, nty = ((uint32_t)g.ny + RBP_OUTY - 1) / RBP_OUTY;
        const uint32_t tiles = ntx * nty;
        const uint32_t items = tiles * (((uint32_t)g.nz + A.zlen - 1) / A.zlen);
        const uint64_t cells = (uint64_t)sz * (uint64_t)g.nz;
        const uint64_t ccells = (uint64_t)csz * (uint64_t)A.cg.nz;
        const RbRsrc Rr = rb_rsrc(rin, cells * PR::esize);
        const RbRsrc Rq = rb_rsrc(upd ? (const void*)A.q : (const void*)rin, cells * 4u);
        const RbRsrc Rd = rb_rsrc(A.dinv, cells * PC::esize);
        const RbRsrc Rx = rb_rsrc(A.bx, cells * PC::esize);
        const RbRsrc Ry = rb_rsrc(A.by, cells * PC::esize);
        const RbRsrc Rz = rb_rsrc(A.bz, cells * PC::esize);
        const RbRsrc Rc = rb_rsrc(MODE == RB_SMOOTH ? (const void*)A.zc : (const void*)A.dinv,
                                  ((MODE == RB_SMOOTH ? ccells : cells) * RbBuf<VT>::size + 3u) & ~(uint64_t)3);
        const RbRsrc Rw = rb_rsrc((const void*)rout, cells * 4u);
        const RbRsrc Ro = rb_rsrc(A.out, (restrict_pass ? ccells : cells) * RbBuf<VT>::size);
        const RbRsrc Rb = rb_rsrc(COARSE && A.base ? (const void*)A.base : (const void*)A.out, cells * PV::esize);
        const int lane = (int)(threadIdx.x & 63u);
        const int lx = lane & (RBP_LX - 1);
        const int ry = (int)(threadIdx.x >> 5);  // tile row: two per wave
        const int ryl = ry > 0 ? ry - 1 : 0, ryh = ry < RBP_NY - 1 ? ry + 1 : RBP_NY - 1;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};  // update: sum r'^2 in [1], sum r' in [3]
        double dot[1] = {0.0};                 // smoother: (r - mean) . t
        for (uint32_t it = vg.b; it < items; it += vg.n) {
            const uint32_t ch = it / tiles, rem = it - ch * tiles;
            const uint32_t ty = rem / ntx, tx = rem - ty * ntx;
            const int xa = (int)tx * RBP_OUTX - 2 + 2 * lx;  // the pair's first column (even)
            const int y = (int)ty * RBP_OUTY - 2 + ry;
            const int za = (int)(ch * A.zlen), zb = min(za + (int)A.zlen, g.nz);
            const int gx = rb_coord(xa, g.nx, perx), gy = rb_coord(y, g.ny, pery);
            const bool lane_ok = gx >= 0 && gy >= 0;
            const uint32_t e0 = lane_ok ? (uint32_t)gx + (uint32_t)gy * (uint32_t)g.nx : 0u;
            // lane byte offsets (past the arrays when the pair is absent: loads read 0)
            const uint32_t vR = lane_ok ? e0 * PR::esize : RB_ABSENT;
            const uint32_t vQ = lane_ok ? e0 * 4u : RB_ABSENT;
            const uint32_t vC = lane_ok ? e0 * PC::esize : RB_ABSENT;
            const uint32_t vV = lane_ok ? e0 * PV::esize : RB_ABSENT;
            const uint32_t eZ = lane_ok ? (uint32_t)(gx >> 1) + (uint32_t)(gy >> 1) * csy : RBPAR_ABSENT;  // parent
            const bool own_x = lx >= 1 && lx <= RBP_LX - 2 && xa < g.nx;
            const bool own_y = ry >= 2 && ry <= RBP_NY - 3 && y < g.ny;
            const bool own = own_x && own_y;
            // ring slots count planes from the chunk's first staged plane za - 2
            float U[3], UB[3], W[3], T0[3], T1[3];
            typename PC::raw BX[3], BY[3], BZ[3], DD[3];
            typename PV::raw BS[3];  // coarse smoother with a base: the base pair of the plane
            float sa = 0.0f;
            auto load = [&](int p) {
                Raw o;
                const int pz = rb_coord(p, g.nz, perz);
                const uint32_t so = pz >= 0 ? sz * (uint32_t)pz : 0u;
                o.r = PR::ld(Rr, vR, so * PR::esize);
                if constexpr (upd) o.q = RbP<float>::ld(Rq, vQ, so * 4u);
                o.d = PC::ld(Rd, vC, so * PC::esize);
                o.bx = PC::ld(Rx, vC, so * PC::esize);
                o.by = PC::ld(Ry, vC, so * PC::esize);
                o.bz = PC::ld(Rz, vC, so * PC::esize);
                if constexpr (MODE == RB_SMOOTH) {
                    o.zc = RbPar<VT>::ld(Rc, eZ + (pz >= 0 ? csz * (uint32_t)(pz >> 1) : 0u));
                    if (COARSE && A.base) o.base = PV::ld(Rb, vV, so * PV::esize);
                }
                return o;
            };
            // loads run two planes ahead of the plane being staged, into a ring slot of their own
            // (a register copy of a load in flight would wait for it)
            Raw RW[3];
            RW[0] = load(za - 2);
            RW[1] = load(za - 1);
            auto step = [&](int p, auto S0c) {
                constexpr int s0 = decltype(S0c)::value, s1 = (s0 + 2) % 3, s2 = (s0 + 1) % 3;
                // slot of plane p - 1 (staged already); issued on every step, also past the chunk's
                // last staged plane, so that the wait for this step's loads never depends on a branch
                RW[s1] = load(p + 2);
                const Raw& cur = RW[s0];
                // red position of the pair on this row at plane p (and the black position at
                // p - 1, the red position at p - 2)
                const bool pr = ((y + p) & 1) != 0;
                // ---- stage 1: plane p
                {
                    const int pzs = rb_coord(p, g.nz, perz);  // uniform
                    const bool pok = pzs >= 0;
                    const typename PC::raw dr = pok ? cur.d : PC::zero();
                    const float d0 = PC::lo(dr), d1 = PC::hi(dr);
                    float rt0, rt1;
                    if constexpr (upd) {
                        const float q0 = __uint_as_float(cur.q.x), q1 = __uint_as_float(cur.q.y);
                        const float rp0 = d0 != 0.0f ? (PR::lo(cur.r) - m) - alpha * q0 : 0.0f;
                        const float rp1 = d1 != 0.0f ? (PR::hi(cur.r) - m) - alpha * q1 : 0.0f;
                        if (own && p >= za && p < zb) {
                            RbP<float>::st(Rw, rp0, rp1, vQ, sz * (uint32_t)p * 4u);
                            acc[1] += (double)rp0 * (double)rp0 + (double)rp1 * (double)rp1;
                            acc[3] += (double)rp0 + (double)rp1;
                        }
                        rt0 = rp0 - mn;
                        rt1 = rp1 - mn;
                    } else {
                        rt0 = PR::lo(cur.r) - m;
                        rt1 = PR::hi(cur.r) - m;
                    }
                    const bool ok = lane_ok && pok;
                    rt0 = ok ? rt0 : 0.0f;
                    rt1 = ok ? rt1 : 0.0f;
                    float u0 = d0 * rt0, u1 = d1 * rt1;
                    if constexpr (MODE == RB_SMOOTH) {
                        if (COARSE && A.base) BS[s0] = cur.base;
                        const float zcv =
                            ok ? ps * RbPar<VT>::f(cur.zc, eZ + (pok ? csz * (uint32_t)(pzs >> 1) : 0u)) : 0.0f;
                        if (pr) u1 += d1 != 0.0f ? zcv : 0.0f;
                        else u0 += d0 != 0.0f ? zcv : 0.0f;
                        T0[s0] = rt0;
                        T1[s0] = rt1;
                    }
                    U[s0] = pr ? u1 : u0;
                    UB[s0] = pr ? u0 : u1;
                    DD[s0] = dr;
                    BX[s0] = pok ? cur.bx : PC::zero();
                    BY[s0] = pok ? cur.by : PC::zero();
                    BZ[s0] = pok ? cur.bz : PC::zero();
                    sU[p & 1][ry][lx] = U[s0];
                    sB[s0][ry][lx] = BY[s0];
                }
                // ---- stage 2: black cells of plane c = p - 1 (position pr)
                if (p >= za) {
                    const int c = p - 1;
                    const float uyl = sU[c & 1][ryl][lx], uyh = sU[c & 1][ryh][lx];
                    const typename PC::raw byh = sB[s1][ryh][lx];
                    const float uc = U[s1];
                    const float up = rb_prev_lane(uc), un = rb_next_lane(uc);
                    const float bxl = PC::lo(BX[s1]), bxh = PC::hi(BX[s1]);
                    const float bxn = rb_next_lane(bxl);
                    float s = pr ? bxh * uc : bxl * up;
                    s += pr ? bxn * un : bxh * uc;
                    s += rbp_sel<CT>(BY[s1], pr) * uyl;
                    s += rbp_sel<CT>(byh, pr) * uyh;
                    s += rbp_sel<CT>(BZ[s1], pr) * U[s2];
                    s += rbp_sel<CT>(BZ[s0], pr) * U[s0];
                    W[s1] = UB[s1] + rbp_sel<CT>(DD[s1], pr) * s;
                    sW[c & 1][ry][lx] = W[s1];
                }
                // ---- stage 3: red cells of plane e = p - 2 (position pr); black values of e - 1
                // in slot s0, e in s2, e + 1 in s1
                const int e = p - 2;
                if (e >= za && e < zb) {
                    const float wyl = sW[e & 1][ryl][lx], wyh = sW[e & 1][ryh][lx];
                    const typename PC::raw byh = sB[s2][ryh][lx];
                    const float wc = W[s2];
                    const float wp = rb_prev_lane(wc), wn = rb_next_lane(wc);
                    const float bxl = PC::lo(BX[s2]), bxh = PC::hi(BX[s2]);
                    const float bxn = rb_next_lane(bxl);
                    float s = pr ? bxh * wc : bxl * wp;
                    s += pr ? bxn * wn : bxh * wc;
                    s += rbp_sel<CT>(BY[s2], pr) * wyl;
                    s += rbp_sel<CT>(byh, pr) * wyh;
                    s += rbp_sel<CT>(BZ[s2], pr) * W[s0];
                    s += rbp_sel<CT>(BZ[s1], pr) * W[s1];
                    const float d = rbp_sel<CT>(DD[s2], pr);
                    if constexpr (restrict_pass) {
                        sa += (own && d != 0.0f) ? s : 0.0f;
                        if ((e & 1) == 1 || e == g.nz - 1) {
                            const float tot = sa + __shfl_xor(sa, 32, 64);  // the row pair (y even, y + 1)
                            if (lane < 32 && own_x && ry >= 2 && ry <= RBP_NY - 3 && y < g.ny)
                                RbBuf<VT>::st(Ro, tot, ((uint32_t)(xa >> 1) + csy * (uint32_t)(y >> 1)) * RbBuf<VT>::size,
                                              csz * (uint32_t)(e >> 1) * RbBuf<VT>::size);
                            sa = 0.0f;
                        }
                    } else {
                        const float rtr = pr ? T1[s2] : T0[s2];
                        const float tr = d * rtr + d * s;
                        float o0 = pr ? wc : tr, o1 = pr ? tr : wc;
                        if (own) {
                            const uint32_t so = sz * (uint32_t)e * PV::esize;
                            if (COARSE && A.base) {
                                o0 = PV::lo(BS[s2]) + A.theta * o0;
                                o1 = PV::hi(BS[s2]) + A.theta * o1;
                            }
                            PV::st(Ro, o0, o1, vV, so);
                            if constexpr (!COARSE) {
                                const typename PC::raw de = DD[s2];
                                if (PC::lo(de) != 0.0f) dot[0] += (double)T0[s2] * (double)o0;
                                if (PC::hi(de) != 0.0f) dot[0] += (double)T1[s2] * (double)o1;
                            }
                        }
                    }
                }
                __syncthreads();
            };
            // whole rounds of three steps (the chunk's zb - za + 4 steps, padded with steps that
            // stage planes past the chunk and store nothing): a loop exit between the steps of a
            // round would make the compiler copy the ring's loads in flight at the back edge
            const int rounds = (zb - za + 6) / 3;
            for (int k = 0, p = za - 2; k < rounds; k++, p += 3) {
                step(p, IC<0>());
                step(p + 1, IC<1>());
                step(p + 2, IC<2>());
            }
        }
        if constexpr (!COARSE) {
            if constexpr (upd) reduce_finish<4, RBP_THREADS>(vg, acc, A.part, sc, 3, FIN_RESTRICT_UPD);
            else if constexpr (MODE == RB_SMOOTH) reduce_finish<1, RBP_THREADS>(vg, dot, A.part, sc, 4, A.mode);
        }
    }
};

// ---------------------------------------------------------------- z-marching operator apply
//
// Pass 1 of the fused iteration with the face mobilities formed in the pass (the brick pass
// K_k_brick_apply_m) on the pair-column layout above: a lane owns a pair of columns, a block of
// 32 x NY lanes a tile of 64 columns by NY rows (the middle 60 x (NY - 2) are outputs) of every
// plane of a z-chunk, marching up z.  Each array is read once per plane of the tile, where a
// brick reads the rows and planes around it again for its neighbours.  Per plane step t: stage
// plane t (search direction f = z + beta p_old, density from the liquid fraction, open bytes)
// into the lane's register rings and the LDS copy of the plane, then apply the operator on plane
// t - 1: x neighbours by DPP lane moves, y neighbours from LDS, z neighbours from the rings.
// Every face mobility is formed from the same two densities and open bit, and the six terms of
// a cell are summed in the same order (x-, x+, y-, y+, z-, z+), as in the brick pass: p and q are
// that pass's bit for bit; only the partial sums of p.Ap and sum q run in another order.  Needs
// nx % 4 == 0 (a pair's open bytes are one half of an aligned dword).
template <class VT>
struct ApzArgs {
    Grid3 g;
    const VT* z;
    float* p0;
    float* p1;
    float* x;
    const float* alpha;  // liquid fraction (two-phase)
    const uint8_t* open8;
    MobP mp;
    float* q;
    double* part;
    Scal* sc;
    int meanfree;
    uint32_t zlen;
};
template <class VT, bool TWO, int NY>
struct K_apz {
    static constexpr int THREADS = RBP_LX * NY;
    static constexpr int MINB = NY == 32 ? 1 : 2;
    static constexpr int OUTY = NY - 2;
    using Args = ApzArgs<VT>;
    using PV = RbP<VT>;
    using PF = RbP<float>;
    struct Raw {  // one plane's loads of a lane (kept raw until staged)
        typename PV::raw z;
        rb_v2u p, a;
        uint32_t o;
    };

    static __device__ void run(const Args& A, const VGrid& vg) {
        Scal* sc = A.sc;
        if (CONVERGED(sc)) return;
        __shared__ float2 sF[2][NY][RBP_LX];    // f of a plane
        __shared__ float2 sD[2][NY][RBP_LX];    // densities of a plane (TWO)
        __shared__ uint32_t sO[2][NY][RBP_LX];  // open bytes of a plane (two per lane)
        const Grid3 g = A.g;
        const MobP mp = A.mp;
        const bool perx = g.per & 1u, pery = g.per & 2u, perz = g.per & 4u;
        const int meanfree = A.meanfree;
        const uint32_t it0 = (uint32_t)sc->iters;
        const float* __restrict__ pold = (it0 & 1u) ? A.p0 : A.p1;
        float* __restrict__ pnew = (it0 & 1u) ? A.p1 : A.p0;
        const float beta = (float)sc->beta;
        const bool pending = sc->pending != 0;
        const float alpha = (float)sc->alpha;
        const float pa = (pending && meanfree) ? pold[sc->anchor] : 0.0f;
        // a solve's first iteration (beta = 0, nothing owed) takes the old search direction as
        // zeros: a refinement step's start leaves it unset
        const bool first = it0 == 0u;
        const uint32_t sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const uint32_t ntx = ((uint32_t)g.nx + RBP_OUTX - 1) / RBP_OUTX, nty = ((uint32_t)g.ny + OUTY - 1) / OUTY;
        const uint32_t tiles = ntx * nty;
        const uint32_t items = tiles * (((uint32_t)g.nz + A.zlen - 1) / A.zlen);
        const uint64_t cells = (uint64_t)sz * (uint64_t)g.nz;
        const RbRsrc Rz = rb_rsrc(A.z, cells * PV::esize);
        const RbRsrc Rp = rb_rsrc(pold, cells * 4u);
        const RbRsrc Ra = rb_rsrc(TWO ? (const void*)A.alpha : (const void*)pold, cells * 4u);
        const RbRsrc Rx = rb_rsrc(A.x, cells * 4u);
        const RbRsrc Ro = rb_rsrc(A.open8, cells);
        const RbRsrc Rn = rb_rsrc(pnew, cells * 4u);
        const RbRsrc Rq = rb_rsrc(A.q, cells * 4u);
        const int lane = (int)(threadIdx.x & 63u);
        const int lx = lane & (RBP_LX - 1);
        const int ry = (int)(threadIdx.x >> 5);  // tile row: two per wave
        const int ryl = ry > 0 ? ry - 1 : 0, ryh = ry < NY - 1 ? ry + 1 : NY - 1;
        double acc[2] = {0.0, 0.0};
        for (uint32_t it = vg.b; it < items; it += vg.n) {
            const uint32_t ch = it / tiles, rem = it - ch * tiles;
            const uint32_t ty = rem / ntx, tx = rem - ty * ntx;
            const int xa = (int)tx * RBP_OUTX - 2 + 2 * lx;  // the pair's first column (even)
            const int y = (int)ty * OUTY - 1 + ry;
            const int za = (int)(ch * A.zlen), zb = min(za + (int)A.zlen, g.nz);
            const int gx = rb_coord(xa, g.nx, perx), gy = rb_coord(y, g.ny, pery);
            const bool lane_ok = gx >= 0 && gy >= 0;
            const bool own = lx >= 1 && lx <= RBP_LX - 2 && xa < g.nx && ry >= 1 && ry <= NY - 2 && y < g.ny;
            const uint32_t e0 = lane_ok ? (uint32_t)gx + (uint32_t)gy * (uint32_t)g.nx : 0u;
            // lane byte offsets (past the arrays when the pair is absent: loads read 0)
            const uint32_t v4 = lane_ok ? e0 * 4u : RB_ABSENT;
            const uint32_t vV = lane_ok ? e0 * PV::esize : RB_ABSENT;
            // open bytes: the aligned dword holding the pair (nx % 4 == 0: the pair is its low or
            // high half); a 16-bit load would make the compiler wait for it at once
            const uint32_t vO = lane_ok ? (e0 & ~3u) : RB_ABSENT, osh = (e0 & 2u) * 8u;
            const uint32_t vX = lane_ok && own && pending ? e0 * 4u : RB_ABSENT;  // x: outputs only
            // ring slots count planes from the chunk's first staged plane za - 1
            // lower-face mobilities (x: the pair's two lower x faces; z) formed once, at staging
            float F0[3], F1[3], D0[3] = {0.0f, 0.0f, 0.0f}, D1[3] = {0.0f, 0.0f, 0.0f}, P0[3], P1[3];
            float WX0[3], WX1[3], WZ0[3], WZ1[3];
            uint32_t OB[3];
            rb_v2u XL[3];  // x of the plane whose operator step is two steps ahead
            auto load = [&](int p) {
                Raw o;
                const int pz = rb_coord(p, g.nz, perz);
                const uint32_t so = pz >= 0 ? sz * (uint32_t)pz : 0u;
                o.z = PV::ld(Rz, vV, so * PV::esize);
                o.p = PF::ld(Rp, v4, so * 4u);
                if constexpr (TWO) o.a = PF::ld(Ra, v4, so * 4u);
                o.o = __builtin_amdgcn_raw_buffer_load_b32(Ro, (int)vO, (int)so, 0);
                return o;
            };
            // loads run two planes ahead of the plane being staged, into a ring slot of their own
            // (a register copy of a load in flight would wait for it)
            Raw RW[3];
            RW[0] = load(za - 1);
            RW[1] = load(za);
            auto step = [&](int t, auto S0c) {
                constexpr int s0 = decltype(S0c)::value, sp = (s0 + 2) % 3, sm = (s0 + 1) % 3;
                // slot of plane t - 1 (staged already); issued on every step, also past the chunk's
                // last staged plane, so that the wait for this step's loads never depends on a branch
                RW[sp] = load(t + 2);
                {  // x of plane t + 1, used by step t + 2 (slot sm held plane t - 2's, used already)
                    const int pz = rb_coord(t + 1, g.nz, perz);
                    XL[sm] = PF::ld(Rx, vX, (pz >= 0 ? sz * (uint32_t)pz : 0u) * 4u);
                }
                // ---- stage plane t
                {
                    const Raw& cur = RW[s0];
                    const bool ok = lane_ok && rb_coord(t, g.nz, perz) >= 0;
                    const float z0 = ok ? PV::lo(cur.z) : 0.0f, z1 = ok ? PV::hi(cur.z) : 0.0f;
                    const float q0 = ok && !first ? __uint_as_float(cur.p.x) : 0.0f;
                    const float q1 = ok && !first ? __uint_as_float(cur.p.y) : 0.0f;
                    F0[s0] = z0 + beta * q0;
                    F1[s0] = z1 + beta * q1;
                    P0[s0] = q0;
                    P1[s0] = q1;
                    const float d0 = TWO ? mob_dens(mp, ok ? __uint_as_float(cur.a.x) : 0.0f) : 0.0f;
                    const float d1 = TWO ? mob_dens(mp, ok ? __uint_as_float(cur.a.y) : 0.0f) : 0.0f;
                    const uint32_t ob = ok ? (cur.o >> osh) & 0xFFFFu : 0u;
                    // ring slot sp still holds plane t - 1 (the plane below)
                    WX0[s0] = mob_face<TWO>(mp, ob & 1u, TWO ? rb_prev_lane(d1) : 0.0f, d0);
                    WX1[s0] = mob_face<TWO>(mp, (ob >> 8) & 1u, d0, d1);
                    WZ0[s0] = mob_face<TWO>(mp, ob & 4u, D0[sp], d0);
                    WZ1[s0] = mob_face<TWO>(mp, (ob >> 8) & 4u, D1[sp], d1);
                    D0[s0] = d0;
                    D1[s0] = d1;
                    OB[s0] = ob;
                    sF[t & 1][ry][lx] = make_float2(F0[s0], F1[s0]);
                    if constexpr (TWO) sD[t & 1][ry][lx] = make_float2(D0[s0], D1[s0]);
                    sO[t & 1][ry][lx] = OB[s0];
                }
                // ---- the operator on plane p = t - 1: rings hold p - 1 in sm, p in sp, p + 1 in s0
                const int p = t - 1;
                if (p >= za && p < zb) {
                    const float2 fyl = sF[p & 1][ryl][lx], fyh = sF[p & 1][ryh][lx];
                    const float2 dyl = TWO ? sD[p & 1][ryl][lx] : make_float2(0.0f, 0.0f);
                    const float2 dyh = TWO ? sD[p & 1][ryh][lx] : make_float2(0.0f, 0.0f);
                    const uint32_t oyh = sO[p & 1][ryh][lx];
                    const float f0 = F0[sp], f1 = F1[sp], d0 = D0[sp], d1 = D1[sp];
                    const uint32_t ob = OB[sp];
                    const float fxm = rb_prev_lane(f1), fxp = rb_next_lane(f0);
                    // column 0's upper x face is column 1's lower one; column 1's upper x face is
                    // the next lane's column 0 lower one; plane p's upper z faces are p + 1's lower
                    const float wx0 = WX0[sp], wx1 = WX1[sp], wx2 = rb_next_lane(WX0[sp]);
                    float a0 = 0.0f;
                    a0 += wx0 * (f0 - fxm);
                    a0 += wx1 * (f0 - f1);
                    a0 += mob_face<TWO>(mp, ob & 2u, dyl.x, d0) * (f0 - fyl.x);
                    a0 += mob_face<TWO>(mp, oyh & 2u, d0, dyh.x) * (f0 - fyh.x);
                    a0 += WZ0[sp] * (f0 - F0[sm]);
                    a0 += WZ0[s0] * (f0 - F0[s0]);
                    float a1 = 0.0f;
                    a1 += wx1 * (f1 - f0);
                    a1 += wx2 * (f1 - fxp);
                    a1 += mob_face<TWO>(mp, (ob >> 8) & 2u, dyl.y, d1) * (f1 - fyl.y);
                    a1 += mob_face<TWO>(mp, (oyh >> 8) & 2u, d1, dyh.y) * (f1 - fyh.y);
                    a1 += WZ1[sp] * (f1 - F1[sm]);
                    a1 += WZ1[s0] * (f1 - F1[s0]);
                    if (own) {
                        const bool act0 = open_active(ob & 0xFFu, meanfree), act1 = open_active((ob >> 8) & 0xFFu, meanfree);
                        const float pn0 = act0 ? f0 : 0.0f, pn1 = act1 ? f1 : 0.0f;
                        acc[0] += (double)pn0 * (double)a0;
                        acc[0] += (double)pn1 * (double)a1;
                        acc[1] += act0 ? (double)a0 : 0.0;
                        acc[1] += act1 ? (double)a1 : 0.0;
                        const uint32_t so = sz * (uint32_t)p * 4u;
                        PF::st(Rn, pn0, pn1, v4, so);
                        PF::st(Rq, a0, a1, v4, so);
                        if (pending) {
                            float x0 = __uint_as_float(XL[sp].x), x1 = __uint_as_float(XL[sp].y);
                            if (act0) x0 = x0 + alpha * (P0[sp] - pa);
                            if (act1) x1 = x1 + alpha * (P1[sp] - pa);
                            PF::st(Rx, x0, x1, v4, so);
                        }
                    }
                }
                __syncthreads();
            };
            // whole rounds of three steps (the chunk's zb - za + 2 steps, padded with steps that
            // stage planes past the chunk and store nothing): a loop exit between the steps of a
            // round would make the compiler copy the ring's loads in flight at the back edge
            const int rounds = (zb - za + 4) / 3;
            for (int k = 0, t = za - 1; k < rounds; k++, t += 3) {
                step(t, IC<0>());
                step(t + 1, IC<1>());
                step(t + 2, IC<2>());
            }
        }
        reduce_finish<2, THREADS>(vg, acc, A.part, sc, 2, FIN_ALPHA);
    }
};

// ---------------------------------------------------------------- tail of the cycle
//
// The smallest red-black levels and the coarsest solve in one launch of one block.  On levels of a
// few thousand cells each pass above is a z-march of a handful of planes, one barrier and one
// memory round trip per plane: 6-8 us per launch, two launches per level and visit.  Here every
// value a level needs (1/diag, residual, the six face couplings of each cell) is loaded once, all
// levels' loads issued together, and kept in registers; a level's values live in LDS, and each
// sweep is one barrier.  Down: pre-smoother (red = u, black sweep), red residuals, restriction;
// the coarsest level's weighted Jacobi sweeps (K_k_coarse); up: prolongation and post-smoother
// (black, then red).  The arithmetic is the passes' (FP32, faces in the order x-, x+, y-, y+, z-,
// z+, the same sums; stored values rounded to the storage type where the passes store them), so
// only the order of a few FP32 roundings may differ.
constexpr int kTailLevels = 2;           // red-black levels in the tail, at most
constexpr uint32_t kTailCells = 4096;    // largest level the tail takes
constexpr uint32_t kTailCoarsest = 64;   // largest coarsest level (every axis at most 4)
constexpr int kTailThreads = 1024;
constexpr int kTailJ = (int)(kTailCells / 2 / kTailThreads);  // pairs per thread on the first tail level

// A thread owns x pairs of cells (x even, x + 1; on a row of odd length the last pair holds one
// cell): one cell of a pair is red and the other black, so each colour's sweep gives every thread
// one cell per pair.
template <class CT, class VT, class AT>
struct K_rb_tail {
    static constexpr int THREADS = kTailThreads;
    static constexpr int MINB = 1;
    struct Args {
        int levels;                        // red-black levels (1..kTailLevels); then the coarsest
        Grid3 g[kTailLevels];
        const CT* d[kTailLevels];          // 1/diag
        const CT* bx[kTailLevels];         // lower-face couplings
        const CT* by[kTailLevels];
        const CT* bz[kTailLevels];
        Grid3 gc;                          // the coarsest level
        const CT* cd;
        const CT* cbx;
        const CT* cby;
        const CT* cbz;
        const CT* react;
        const VT* r0;                      // residual of the first tail level
        VT* t0;                            // its correction (written; deeper levels stay in LDS)
        Scal* sc;
    };

    // neighbour coordinate along one axis, or -1 past a closed end
    static __device__ __forceinline__ int nbr(int v, int n, bool per) {
        if (v < 0) return per ? v + n : -1;
        if (v >= n) return per ? v - n : -1;
        return v;
    }
    // the six neighbours of cell i = (x, y, z) in face order (the cell itself past a closed end)
    static __device__ __forceinline__ void nbrs(const Grid3& g, int x, int y, int z, uint32_t i, uint32_t (&nb)[6],
                                                bool (&ok)[6]) {
        const bool px = g.per & 1u, py = g.per & 2u, pz = g.per & 4u;
        const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
        const int c[6] = {nbr(x - 1, g.nx, px), nbr(x + 1, g.nx, px), nbr(y - 1, g.ny, py),
                          nbr(y + 1, g.ny, py), nbr(z - 1, g.nz, pz), nbr(z + 1, g.nz, pz)};
        const uint32_t base[3] = {i - (uint32_t)x, i - (uint32_t)y * sy, i - (uint32_t)z * sz};
        const uint32_t str[3] = {1u, sy, sz};
#pragma unroll
        for (int f = 0; f < 6; f++) {
            ok[f] = c[f] >= 0;
            nb[f] = ok[f] ? base[f >> 1] + (uint32_t)c[f] * str[f >> 1] : i;
        }
    }
    // the six face couplings: lower faces the cell's own, upper faces the neighbour's lower ones
    static __device__ __forceinline__ void couplings(const CT* bx, const CT* by, const CT* bz, uint32_t i,
                                                     const uint32_t (&nb)[6], const bool (&ok)[6], float (&b)[6]) {
        b[0] = ok[0] ? ldf(bx[i]) : 0.0f;
        b[1] = ok[1] ? ldf(bx[nb[1]]) : 0.0f;
        b[2] = ok[2] ? ldf(by[i]) : 0.0f;
        b[3] = ok[3] ? ldf(by[nb[3]]) : 0.0f;
        b[4] = ok[4] ? ldf(bz[i]) : 0.0f;
        b[5] = ok[5] ? ldf(bz[nb[5]]) : 0.0f;
    }
    static __device__ __forceinline__ uint32_t pack(int x, int y, int z) {
        return (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
    }
    // x pairs of a level: ceil(nx / 2) per row (the last one of an odd row holds one cell)
    static __device__ __forceinline__ uint32_t pairs_of(const Grid3& g) {
        return (uint32_t)((g.nx + 1) / 2) * (uint32_t)g.ny * (uint32_t)g.nz;
    }
    // pair p: its first cell (x even, y, z), the cell's index, whether the second cell exists
    static __device__ __forceinline__ uint32_t pair_cell(const Grid3& g, uint32_t p, int& x, int& y, int& z, bool& has2) {
        const uint32_t ppr = (uint32_t)((g.nx + 1) / 2), row = p / ppr;
        x = (int)(2u * (p - row * ppr));
        y = (int)(row % (uint32_t)g.ny);
        z = (int)(row / (uint32_t)g.ny);
        has2 = x + 1 < g.nx;
        return row * (uint32_t)g.nx + (uint32_t)x;
    }
    // sum over the faces of cell (x0 + h, y, z) of coupling * v(neighbour), in face order; xyz: the
    // pair's first cell, i its index
    static __device__ __forceinline__ float fsum(const Grid3& g, uint32_t xyz, int h, uint32_t i, const float (&b)[6],
                                                 const float* v) {
        uint32_t nb[6];
        bool ok[6];
        nbrs(g, (int)(xyz & 1023u) + h, (int)((xyz >> 10) & 1023u), (int)((xyz >> 20) & 1023u), i + (uint32_t)h, nb, ok);
        float s = b[0] * v[nb[0]];
#pragma unroll
        for (int f = 1; f < 6; f++) s += b[f] * v[nb[f]];
        return s;
    }
    static __device__ __forceinline__ float rnd(float v) { return ldf(cvt<VT>(v)); }
    // the coarse cell c's restricted residual: its red children, (row y0: z0 + z1) + (row y1: ...)
    static __device__ __forceinline__ float restrict_at(const Grid3& g, const Grid3& cg, uint32_t c, const float* S) {
        int cx, cy, cz_;
        decode(cg, c, cx, cy, cz_);
        float rows[2] = {0.0f, 0.0f};
#pragma unroll
        for (int dy = 0; dy < 2; dy++) {
            const int y = 2 * cy + dy;
            if (y >= g.ny) continue;
            float sa = 0.0f;
#pragma unroll
            for (int dz = 0; dz < 2; dz++) {
                const int z = 2 * cz_ + dz;
                if (z >= g.nz) continue;
                const int x = 2 * cx + ((y + z) & 1);
                if (x >= g.nx) continue;
                sa += S[(uint32_t)x + (uint32_t)g.nx * ((uint32_t)y + (uint32_t)g.ny * (uint32_t)z)];
            }
            rows[dy] = sa;
        }
        return rnd(rows[0] + rows[1]);
    }

    static __device__ void run(const Args& A, const VGrid& vg) {
        (void)vg;
        Scal* sc = A.sc;
        if (CONVERGED(sc)) return;
        __shared__ float S[kTailCells];            // the current level's values
        __shared__ float TC[kTailCells / 8];       // the level below's correction (prolonged up)
        __shared__ AT cz[kTailCoarsest];
        const float ps = (float)sc->pscale;
        const uint32_t tid = threadIdx.x;
        const int L = A.levels;
        // ---- every level's values, loaded together and kept by colour (red, black) of each owned
        // pair: 1/diag, face couplings, and the first level's residual (deeper residuals come from
        // the restriction)
        float DR[kTailLevels][kTailJ], DB[kTailLevels][kTailJ], RR[kTailLevels][kTailJ], RB[kTailLevels][kTailJ];
        float BR[kTailLevels][kTailJ][6], BB[kTailLevels][kTailJ][6];
        uint32_t XYZ[kTailLevels][kTailJ];  // the pair's first cell; bit 31: the red cell is the second
        uint32_t I0[kTailLevels][kTailJ];   // the first cell's index; bit 31: the second cell exists
#pragma unroll
        for (int k = 0; k < kTailLevels; k++)
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {
                XYZ[k][j] = I0[k][j] = 0u;
                DR[k][j] = DB[k][j] = RR[k][j] = RB[k][j] = 0.0f;
#pragma unroll
                for (int f = 0; f < 6; f++) BR[k][j][f] = BB[k][j][f] = 0.0f;
                if (k >= L || (k > 0 && j > 0)) continue;
                const Grid3 g = A.g[k];
                const uint32_t p = tid + (uint32_t)j * kTailThreads;
                if (p >= pairs_of(g)) continue;
                int x, y, z;
                bool has2;
                const uint32_t i0 = pair_cell(g, p, x, y, z, has2);
                const bool hr = ((y + z) & 1) != 0;
                XYZ[k][j] = pack(x, y, z) | (hr ? 0x80000000u : 0u);
                I0[k][j] = i0 | (has2 ? 0x80000000u : 0u);
                float d[2] = {0.0f, 0.0f}, r[2] = {0.0f, 0.0f}, b[2][6];
#pragma unroll
                for (int h = 0; h < 2; h++) {
#pragma unroll
                    for (int f = 0; f < 6; f++) b[h][f] = 0.0f;
                    if (h == 1 && !has2) continue;
                    const uint32_t i = i0 + (uint32_t)h;
                    d[h] = ldf(A.d[k][i]);
                    if (k == 0) r[h] = ldf(A.r0[i]);
                    uint32_t nb[6];
                    bool ok[6];
                    nbrs(g, x + h, y, z, i, nb, ok);
                    couplings(A.bx[k], A.by[k], A.bz[k], i, nb, ok, b[h]);
                }
                DR[k][j] = hr ? d[1] : d[0];
                DB[k][j] = hr ? d[0] : d[1];
                RR[k][j] = hr ? r[1] : r[0];
                RB[k][j] = hr ? r[0] : r[1];
#pragma unroll
                for (int f = 0; f < 6; f++) {
                    BR[k][j][f] = hr ? b[1][f] : b[0][f];
                    BB[k][j][f] = hr ? b[0][f] : b[1][f];
                }
            }
        // the coarsest level: K_k_coarse's values, one cell per thread of the first wave
        const Grid3 gc = A.gc;
        const bool in_c = tid < gc.n;
        AT crhs = 0, cdi = 0, crc = 0;
        AT cb[6] = {0, 0, 0, 0, 0, 0};
        uint32_t cnb[6] = {0, 0, 0, 0, 0, 0};
        if (in_c) {
            int x, y, z;
            decode(gc, tid, x, y, z);
            bool ok[6];
            float b[6];
            nbrs(gc, x, y, z, tid, cnb, ok);
            couplings(A.cbx, A.cby, A.cbz, tid, cnb, ok, b);
            cdi = (AT)ldf(A.cd[tid]);
            crc = (AT)ldf(A.react[tid]);
#pragma unroll
            for (int f = 0; f < 6; f++) cb[f] = (AT)b[f];
        }
        auto red_at = [](uint32_t xyz) { return (uint32_t)(xyz >> 31); };  // offset of the red cell
        auto first = [](uint32_t i0) { return i0 & 0x7FFFFFFFu; };
        // the cell at offset h of the pair exists (the second one may not, at the end of an odd row)
        auto has = [](uint32_t i0, uint32_t h) { return h == 0u || (i0 >> 31) != 0u; };
        // ---- down: pre-smoother, red residuals, restriction
#pragma unroll
        for (int k = 0; k < kTailLevels; k++) {
            if (k >= L) break;
            const Grid3 g = A.g[k];
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {
                if (!(k == 0 || j == 0) || tid + (uint32_t)j * kTailThreads >= pairs_of(g)) continue;
                const uint32_t i0 = first(I0[k][j]), hr = red_at(XYZ[k][j]);
                if (has(I0[k][j], hr)) S[i0 + hr] = DR[k][j] * RR[k][j];
                if (has(I0[k][j], 1u - hr)) S[i0 + 1u - hr] = DB[k][j] * RB[k][j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {  // black sweep
                if (!(k == 0 || j == 0) || tid + (uint32_t)j * kTailThreads >= pairs_of(g)) continue;
                const uint32_t i0 = first(I0[k][j]), hb = 1u - red_at(XYZ[k][j]);
                if (!has(I0[k][j], hb)) continue;
                S[i0 + hb] = DB[k][j] * RB[k][j] + DB[k][j] * fsum(g, XYZ[k][j] & 0x7FFFFFFFu, (int)hb, i0, BB[k][j], S);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {  // red residuals (black residuals are zero)
                if (!(k == 0 || j == 0) || tid + (uint32_t)j * kTailThreads >= pairs_of(g)) continue;
                const uint32_t i0 = first(I0[k][j]), hr = red_at(XYZ[k][j]);
                if (!has(I0[k][j], hr)) continue;
                const float s = fsum(g, XYZ[k][j] & 0x7FFFFFFFu, (int)hr, i0, BR[k][j], S);
                S[i0 + hr] = DR[k][j] != 0.0f ? s : 0.0f;
            }
            __syncthreads();
            // restriction into the next level's owned pairs, or the coarsest level's cells
            if (k + 1 < L) {
                const int kn = k + 1 < kTailLevels ? k + 1 : 0;
                const Grid3 cg = A.g[kn];
                if (tid < pairs_of(cg)) {
                    const uint32_t c0 = first(I0[kn][0]);
                    const float r0 = restrict_at(g, cg, c0, S);
                    const float r1 = has(I0[kn][0], 1u) ? restrict_at(g, cg, c0 + 1, S) : 0.0f;
                    const bool hr = red_at(XYZ[kn][0]) != 0u;
                    RR[kn][0] = hr ? r1 : r0;
                    RB[kn][0] = hr ? r0 : r1;
                }
            } else if (in_c) {
                const float rc = restrict_at(g, gc, tid, S);
                crhs = cdi != (AT)0 ? (AT)rc : (AT)0;
            }
            __syncthreads();
        }
        // ---- the coarsest level: eight weighted Jacobi sweeps from zero, by the first wave alone
        // (at most 64 cells): a wave's LDS accesses complete in order, so between its sweeps the
        // lanes need only a wave-scope fence, not a block barrier
        static_assert(kTailCoarsest <= 64, "the coarsest level fits one wave");
        if (tid < 64) {
            auto wave_sync = [] {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            };
            cz[tid] = 0;
            wave_sync();
            const AT omega = (AT)sc->omega;
            for (int step = 0; step < 8; step++) {
                AT next = 0;
                if (in_c) {
                    const AT center = cz[tid];
                    if (cdi != (AT)0) {
                        AT a = crc * center;
#pragma unroll
                        for (int f = 0; f < 6; f++) a += cb[f] * (center - cz[cnb[f]]);
                        next = center + omega * (crhs - a) * cdi;
                    }
                }
                wave_sync();
                if (in_c) cz[tid] = next;
                wave_sync();
            }
            if (in_c) TC[tid] = rnd((float)cz[tid]);
        }
        __syncthreads();
        // ---- up: prolongation, post-smoother (black, then red)
#pragma unroll
        for (int k = kTailLevels - 1; k >= 0; k--) {
            if (k >= L) continue;
            const Grid3 g = A.g[k];
            const Grid3 cg = k + 1 < L ? A.g[k + 1 < kTailLevels ? k + 1 : 0] : gc;
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {
                if (!(k == 0 || j == 0) || tid + (uint32_t)j * kTailThreads >= pairs_of(g)) continue;
                const uint32_t i0 = first(I0[k][j]);
                const uint32_t xyz = XYZ[k][j];
                const uint32_t hr = red_at(xyz);
                const int x = (int)(xyz & 1023u) + (int)hr, y = (int)((xyz >> 10) & 1023u), z = (int)((xyz >> 20) & 1023u);
                const uint32_t pc = (uint32_t)(x >> 1) + (uint32_t)cg.nx * ((uint32_t)(y >> 1) + (uint32_t)cg.ny * (uint32_t)(z >> 1));
                const float zcv = ps * TC[pc];
                float u = DR[k][j] * RR[k][j];
                u += DR[k][j] != 0.0f ? zcv : 0.0f;
                if (has(I0[k][j], hr)) S[i0 + hr] = u;
                if (has(I0[k][j], 1u - hr)) S[i0 + 1u - hr] = DB[k][j] * RB[k][j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {  // black sweep
                if (!(k == 0 || j == 0) || tid + (uint32_t)j * kTailThreads >= pairs_of(g)) continue;
                const uint32_t i0 = first(I0[k][j]), hb = 1u - red_at(XYZ[k][j]);
                if (!has(I0[k][j], hb)) continue;
                S[i0 + hb] = DB[k][j] * RB[k][j] + DB[k][j] * fsum(g, XYZ[k][j] & 0x7FFFFFFFu, (int)hb, i0, BB[k][j], S);
            }
            __syncthreads();
            float ored[kTailJ], oblack[kTailJ];
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {  // red sweep; the level's correction
                ored[j] = oblack[j] = 0.0f;
                if (!(k == 0 || j == 0) || tid + (uint32_t)j * kTailThreads >= pairs_of(g)) continue;
                const uint32_t i0 = first(I0[k][j]), hr = red_at(XYZ[k][j]);
                if (has(I0[k][j], hr))
                    ored[j] = DR[k][j] * RR[k][j] + DR[k][j] * fsum(g, XYZ[k][j] & 0x7FFFFFFFu, (int)hr, i0, BR[k][j], S);
                if (has(I0[k][j], 1u - hr)) oblack[j] = S[i0 + 1u - hr];
            }
            // TC still holds the level below's correction until every thread has read it (above)
#pragma unroll
            for (int j = 0; j < kTailJ; j++) {
                if (!(k == 0 || j == 0) || tid + (uint32_t)j * kTailThreads >= pairs_of(g)) continue;
                const uint32_t i0 = first(I0[k][j]), hr = red_at(XYZ[k][j]);
                const bool hasr = has(I0[k][j], hr), hasb = has(I0[k][j], 1u - hr);
                if (k == 0) {
                    if (hasr) A.t0[i0 + hr] = cvt<VT>(ored[j]);
                    if (hasb) A.t0[i0 + 1u - hr] = cvt<VT>(oblack[j]);
                } else {
                    if (hasr) TC[i0 + hr] = rnd(ored[j]);
                    if (hasb) TC[i0 + 1u - hr] = rnd(oblack[j]);
                }
            }
            __syncthreads();
        }
    }
};

// The first level of the tail (L when there is none): the largest level from 2 on of at most
// kTailCells cells such that it and every level below it but the coarsest are red-black levels,
// at most kTailLevels of them, and the coarsest level fits the tail's solve.
int rb_tail_level(int L, const LevelDev* lv) {
    static const bool on = [] {
        const char* e = getenv("MHDP_RB_TAIL");
        return !(e && e[0] == '0');
    }();
    if (!on || L < 4 || lv[L - 1].g.n > kTailCoarsest) return L;
    // the first tail level's pairs fill at most kTailJ per thread (the deeper levels' at most 512
    // cells, one pair per thread)
    auto fits = [](const Grid3& g) {
        const uint64_t pairs = (uint64_t)((g.nx + 1) / 2) * (uint64_t)g.ny * (uint64_t)g.nz;
        return g.n <= kTailCells && pairs <= (uint64_t)kTailJ * kTailThreads && g.nx < 1024 && g.ny < 1024 &&
               g.nz < 1024;
    };
    int T = L - 1;
    while (T - 1 >= 2 && fits(lv[T - 1].g) && rb_level_ok(lv[T - 1].g) && L - 1 - (T - 1) <= kTailLevels &&
           (T - 1 == L - 2 || lv[T].g.n <= kTailCells / 8))
        T--;
    return T < L - 1 ? T : L;
}

template <class CT, class VT, class AT>
static void rb_tail_go(hipStream_t s, bool f16, int L, LevelDev* lv, int T, Scal* sc) {
    typename K_rb_tail<CT, VT, AT>::Args a{};
    a.levels = L - 1 - T;
    auto arr = [&](const LevelDev& l, int which) {
        void* const a16[5] = {l.bx16, l.by16, l.bz16, l.react16, l.dinv16};
        float* const a32[5] = {l.bx, l.by, l.bz, l.react, l.dinv};
        return f16 ? (const CT*)a16[which] : (const CT*)a32[which];
    };
    for (int k = 0; k < a.levels; k++) {
        const LevelDev& l = lv[T + k];
        a.g[k] = l.g;
        a.bx[k] = arr(l, 0);
        a.by[k] = arr(l, 1);
        a.bz[k] = arr(l, 2);
        a.d[k] = arr(l, 4);
    }
    const LevelDev& c = lv[L - 1];
    a.gc = c.g;
    a.cbx = arr(c, 0);
    a.cby = arr(c, 1);
    a.cbz = arr(c, 2);
    a.react = arr(c, 3);
    a.cd = arr(c, 4);
    a.r0 = (const VT*)lv[T].r;
    a.t0 = (VT*)lv[T].t;
    a.sc = sc;
    go<K_rb_tail<CT, VT, AT>>(s, 1, a);
}

void launch_rb_tail(hipStream_t s, int store, bool coarse_f64, int L, LevelDev* lv, int T, Scal* sc) {
    using T16 = __hip_bfloat16;
    if (store == STORE_BF16) {
        if (coarse_f64) rb_tail_go<T16, T16, double>(s, true, L, lv, T, sc);
        else rb_tail_go<T16, T16, float>(s, true, L, lv, T, sc);
    } else {
        if (coarse_f64) rb_tail_go<float, float, double>(s, false, L, lv, T, sc);
        else rb_tail_go<float, float, float>(s, false, L, lv, T, sc);
    }
}

// ---------------------------------------------------------------- launchers

bool rb_level_ok(const Grid3& g) {
    // buffer addressing: every array of the level below 2 GiB (4-byte elements)
    if ((uint64_t)g.nx * (uint64_t)g.ny * (uint64_t)g.nz * 4u >= (1ull << 31)) return false;
    for (int a = 0; a < 3; a++) {
        const int n = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
        if (((g.per >> a) & 1u) && (n & 1)) return false;
    }
    return true;
}

constexpr int RBW_ROWS = 4;  // rows of a wave strip

// Planes per z-chunk of the wave-strip kernel: some eight thousand work items (a few per resident
// wave) where the level is large enough; even, 8..64.
static uint32_t rbw_zlen(const Grid3& g) {
    const uint64_t tiles = (uint64_t)((g.nx + RBW_OUT - 1) / RBW_OUT) * (uint64_t)((g.ny + RBW_ROWS - 1) / RBW_ROWS);
    uint64_t z = (uint64_t)g.nz * tiles / 8192u;
    z &= ~(uint64_t)1;
    return (uint32_t)(z < 8 ? 8 : (z > 64 ? 64 : z));
}

// Planes per z-chunk of the pair-column kernel.  A coarse level's time is its blocks' serial
// marches: the even chunk length in 4..64 with the fewest plane steps per block (rounds of work
// items times the chunk plus the two planes staged below and above it).  The fine level streams
// HBM: about two thousand items (even, 8..64), which keeps the blocks in step and the chunk ends'
// re-read planes few.
static uint32_t rbp_zlen(const Grid3& g, int outy, uint32_t blocks, bool coarse) {
    const uint64_t tiles = (uint64_t)((g.nx + RBP_OUTX - 1) / RBP_OUTX) * (uint64_t)((g.ny + outy - 1) / outy);
    if (!coarse && outy < 28) {  // fine levels in 16-row tiles: 8 planes (their grids are small)
        uint64_t z = (uint64_t)g.nz * tiles * 3u / 8192u;
        z = (z + 1) / 6 * 6 + 2;
        return (uint32_t)(z < 8 ? 8 : (z > 62 ? 62 : z));
    }
    if (!coarse) {
        // fine levels in 32-row tiles: the fewest plane steps per resident block among chunks of
        // 8..32 planes with z + 4 a whole number of three-step rounds (256^3: 26 planes, 11.45 ms per
        // nozzle projection against 12.31 ms at 8 and 12.15 ms at 32; 512^3: 32 planes, 41.3 ms
        // against 41.4 at 26, 41.6 at 44 and 42.0 ms at 62)
        uint32_t best = 8;
        uint64_t best_steps = ~(uint64_t)0;
        for (uint32_t z = 8; z <= 32; z += 6) {
            const uint64_t items = tiles * (((uint64_t)g.nz + z - 1) / z);
            const uint64_t steps = (items + blocks - 1) / blocks * (z + 4);
            if (steps < best_steps) best_steps = steps, best = z;
        }
        return best;
    }
    uint32_t best = 4;
    uint64_t best_steps = ~(uint64_t)0;
    for (uint32_t z = 4; z <= 64; z += 2) {
        const uint64_t items = tiles * (((uint64_t)g.nz + z - 1) / z);
        const uint64_t steps = (items + blocks - 1) / blocks * ((z + 6) / 3 * 3);
        if (steps < best_steps) best_steps = steps, best = z;
        if (z >= (uint32_t)g.nz) break;
    }
    return best;
}

// The pair-column kernel takes levels whose rows hold whole pairs (nx even); MHDP_RB_KERNEL=strip
// keeps the wave-strip kernel everywhere.
static bool rbp_enabled() {
    static const bool on = [] {
        const char* e = getenv("MHDP_RB_KERNEL");
        return !(e && e[0] == 's');
    }();
    return on;
}

// Coarse levels of fewer cells than this run the 16-row tiles, and so do fine levels of fewer than
// kRbpSmallFine cells (with the fine chunk rule): two blocks per CU hide the latency of their short
// marches (64^3: 1.58 -> 1.43 ms per mixed projection; the 64-problem batch of 32^3..96^3 grids
// 20.2 -> 18.2 ms; at 128^3 the two tile heights take the same time).  A 256^3 coarse level (the
// first of a 512^3 grid) runs the 32-row tiles: 50-63 us per pass against 53-88 us, 41.7 -> 41.4 ms
// per 512^3 projection.
constexpr uint32_t kRbpSmallCells = 1u << 24;
constexpr uint32_t kRbpSmallFine = 1u << 22;

template <class KP, bool COARSE, class A>
static void rbp_launch(hipStream_t s, A a) {
    // the resident blocks (one or two per CU): they march in step, so the halo rows and columns
    // one tile shares with the next are still in L2
    const uint32_t cap = 256u * (uint32_t)KP::MINB;
    a.zlen = rbp_zlen(a.g, KP::RBP_OUTY, cap, COARSE);
    const uint32_t tiles = (uint32_t)((a.g.nx + RBP_OUTX - 1) / RBP_OUTX) * (uint32_t)((a.g.ny + KP::RBP_OUTY - 1) / KP::RBP_OUTY);
    const uint32_t items = tiles * (((uint32_t)a.g.nz + a.zlen - 1) / a.zlen);
    uint32_t blocks = items < cap ? items : cap;
    if (blocks >= 64) blocks = (blocks + MHDP_XCDS - 1) / MHDP_XCDS * MHDP_XCDS;
    go<KP>(s, blocks, a);
}

// The z-marching apply replaces the brick apply with MHDP_APPLY_MARCH=1 (grids with nx % 4 == 0).
// It is not the default: at 512^3 it takes 880 us per launch against the brick pass's 847 us
// (it forms four face mobilities per cell with IEEE divisions, as the brick does, but with a
// barrier per plane and less independent work per lane to hide them).
static bool apz_enabled() {
    static const bool on = [] {
        const char* e = getenv("MHDP_APPLY_MARCH");
        return e && e[0] && e[0] != '0';
    }();
    return on;
}
// z-chunk of the apply: the fewest plane steps per resident block (a chunk of z planes costs its
// z + 2 steps with the two halo planes, rounded up to whole rounds of three), among chunks of at
// most 12 planes (longer ones ran slower at 512^3: 16 planes 916 us, 64 planes 1163 us, against
// 880 us at 10).
static uint32_t apz_zlen(const Grid3& g, int outy, uint32_t blocks) {
    const uint64_t tiles = (uint64_t)((g.nx + RBP_OUTX - 1) / RBP_OUTX) * (uint64_t)((g.ny + outy - 1) / outy);
    uint32_t best = 4;
    uint64_t best_steps = ~(uint64_t)0;
    for (uint32_t z = 4; z <= 12; z++) {
        const uint64_t items = tiles * (((uint64_t)g.nz + z - 1) / z);
        const uint64_t steps = (items + blocks - 1) / blocks * ((z + 4) / 3 * 3);
        if (steps < best_steps) best_steps = steps, best = z;
        if (z >= (uint32_t)g.nz) break;
    }
    return best;
}
template <class KP>
static void apz_launch(hipStream_t s, typename KP::Args a) {
    const uint32_t cap = 256u * (uint32_t)KP::MINB;  // the resident blocks
    a.zlen = apz_zlen(a.g, KP::OUTY, cap);
    const uint32_t tiles = (uint32_t)((a.g.nx + RBP_OUTX - 1) / RBP_OUTX) * (uint32_t)((a.g.ny + KP::OUTY - 1) / KP::OUTY);
    const uint32_t items = tiles * (((uint32_t)a.g.nz + a.zlen - 1) / a.zlen);
    uint32_t blocks = items < cap ? items : cap;
    if (blocks >= 64) blocks = (blocks + MHDP_XCDS - 1) / MHDP_XCDS * MHDP_XCDS;
    go<KP>(s, blocks, a);
}
template <class VT, bool TWO>
static void apz_go(hipStream_t s, const ApzArgs<VT>& a) {
    if (a.g.n < kRbpSmallFine) apz_launch<K_apz<VT, TWO, 16>>(s, a);
    else apz_launch<K_apz<VT, TWO, 32>>(s, a);
}

template <class CT, class VT, class RT, int MODE, bool COARSE>
static void rb_go(hipStream_t s, const typename K_rb<CT, VT, RT, MODE, COARSE, RBW_ROWS>::Args& a0) {
    using K = K_rb<CT, VT, RT, MODE, COARSE, RBW_ROWS>;
    typename K::Args a = a0;
    if (rbp_enabled() && (a.g.nx & 1) == 0) {
        if (a.g.n < (COARSE ? kRbpSmallCells : kRbpSmallFine))
            rbp_launch<K_rbp<CT, VT, RT, MODE, COARSE, 16>, COARSE>(s, a);
        else
            rbp_launch<K_rbp<CT, VT, RT, MODE, COARSE, 32>, COARSE>(s, a);
        return;
    }
    a.zlen = rbw_zlen(a.g);
    const uint32_t tiles = (uint32_t)((a.g.nx + RBW_OUT - 1) / RBW_OUT) * (uint32_t)((a.g.ny + RBW_ROWS - 1) / RBW_ROWS);
    const uint32_t items = tiles * (((uint32_t)a.g.nz + a.zlen - 1) / a.zlen);
    go<K>(s, grid_blocks(items * 64u, launch_tuning().march_cap), a);  // one wave per item
}

void launch_rb_restrict(hipStream_t s, int store, bool coarse, const Grid3& cg, const Grid3& g, const void* r,
                        const void* r2, float* w0, float* w1, const float* q, const void* dinv, const void* bx,
                        const void* by, const void* bz, void* rc, double* part, Scal* sc) {
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        if (coarse)
            rb_go<T, T, T, RB_RESTRICT, true>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                    (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, nullptr,
                                                    (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else if (q)
            rb_go<T, T, float, RB_RESTRICT_UPDATE, false>(
                s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q, (const T*)dinv, (const T*)bx, (const T*)by,
                    (const T*)bz, nullptr, (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_RESTRICT, false>(s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q,
                                                         (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz,
                                                         nullptr, (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
    } else {
        using T = float;
        if (coarse)
            rb_go<T, T, T, RB_RESTRICT, true>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                   (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz, nullptr,
                                                   (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else if (q)
            rb_go<T, T, float, RB_RESTRICT_UPDATE, false>(
                s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q, (const T*)dinv, (const T*)bx, (const T*)by,
                    (const T*)bz, nullptr, (T*)rc, nullptr, 0.0f, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_RESTRICT, false>(s, {g, cg, (const float*)r, (const float*)r2, w0, w1, q,
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
            rb_go<T, T, T, RB_SMOOTH, true>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                  (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz,
                                                  (const T*)zc, (T*)t, (const T*)base, theta, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_SMOOTH, false>(s, {g, cg, (const float*)r, (const float*)r2, nullptr, nullptr,
                                                       nullptr, (const T*)dinv, (const T*)bx, (const T*)by,
                                                       (const T*)bz, (const T*)zc, (T*)t, nullptr, 0.0f, part, sc,
                                                       mode, use_next ? 1 : 0, 0});
    } else {
        using T = float;
        if (coarse)
            rb_go<T, T, T, RB_SMOOTH, true>(s, {g, cg, (const T*)r, (const T*)r, nullptr, nullptr, nullptr,
                                                 (const T*)dinv, (const T*)bx, (const T*)by, (const T*)bz,
                                                 (const T*)zc, (T*)t, (const T*)base, theta, part, sc, 0, 0, 0});
        else
            rb_go<T, T, float, RB_SMOOTH, false>(s, {g, cg, (const float*)r, (const float*)r2, nullptr, nullptr,
                                                      nullptr, (const T*)dinv, (const T*)bx, (const T*)by,
                                                      (const T*)bz, (const T*)zc, (T*)t, nullptr, 0.0f, part, sc,
                                                      mode, use_next ? 1 : 0, 0});
    }
}

// launch_brick_apply_m on the z-marching kernel; false (nothing launched) where it does not apply.
bool launch_apply_march(hipStream_t s, const Grid3& g, int store, const void* z, float* p0, float* p1, float* x,
                        const float* alpha, const uint8_t* open8, float rho_l, float rho_air, float* q, double* part,
                        Scal* sc, int meanfree) {
    if ((g.nx & 3) || !apz_enabled()) return false;
    const MobP mp{rho_l, rho_air};
    if (store == STORE_BF16) {
        using T = __hip_bfloat16;
        if (alpha) apz_go<T, true>(s, {g, (const T*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree, 0});
        else apz_go<T, false>(s, {g, (const T*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree, 0});
    } else {
        if (alpha) apz_go<float, true>(s, {g, (const float*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree, 0});
        else apz_go<float, false>(s, {g, (const float*)z, p0, p1, x, alpha, open8, mp, q, part, sc, meanfree, 0});
    }
    return true;
}
