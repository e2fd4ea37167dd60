// The momentum step either side of the projection (solver.wgsl:127-163, mod.rs:1285-1303):
//   predict  semi-Lagrangian RK2 advection of the face velocities (departure point from the
//            velocity sampled at the face centre and at the half-step midpoint, trilinear samples)
//   stress   the two-phase viscous stress diffusion (viscosity.wgsl, stabilized.wgsl): explicit
//            substeps or damped RKC2 stages of the stress divergence
//   viscous  explicit face Laplacian with no-slip reflection at closed tangential neighbours,
//            body + EM (+ capillary, two-phase) acceleration, times the face mobility.
// Built with -ffp-contract=off: every FP32 expression keeps the reference's operation order, so
// the results equal the oracle's bit for bit.
//
// Faces are the three SoA planes (plane a has dims n + e_a); each kernel is instantiated per
// plane (the axis is a template argument).  Which faces are open depends only on the mask, so it
// is kept per cell as one byte (bit a: the cell's lower face on axis a is open; bit 3: the cell is
// fluid), rebuilt when the cells change; an open test is then one byte load instead of two mask
// loads.  predict first writes the open-face copy of its input (closed faces 0), so each of its 56
// samples per face is one load (periodic wrap and range check stay).
#include <cstdlib>

#include "launch.h"
#include "mhdp_kernels.h"

namespace {

struct Planes {
    int n[3];
    uint32_t per;
    uint32_t f0, f1;  // sizes of the x and y planes
};

Planes planes_of(const Grid3& g) {
    Planes p;
    p.n[0] = g.nx;
    p.n[1] = g.ny;
    p.n[2] = g.nz;
    p.per = g.per;
    p.f0 = ((uint32_t)g.nx + 1) * (uint32_t)g.ny * (uint32_t)g.nz;
    p.f1 = (uint32_t)g.nx * ((uint32_t)g.ny + 1) * (uint32_t)g.nz;
    return p;
}
uint32_t plane_size(const Planes& p, int a) {
    uint32_t s = 1;
    for (int b = 0; b < 3; b++) s *= (uint32_t)p.n[b] + (b == a ? 1u : 0u);
    return s;
}

// wrap_periodic (src/periodic.wgsl:5-10): unsigned Euclidean wrap of any integer
__device__ __forceinline__ int wrap_any(int v, int size) {
    if (v >= 0) return (int)((uint32_t)v % (uint32_t)size);
    const uint32_t magnitude = (uint32_t)(-(v + 1)) + 1u;
    const uint32_t rem = magnitude % (uint32_t)size;
    return rem == 0u ? 0 : (int)((uint32_t)size - rem);
}
// the same, without a division (and without branches) for values within one period of the range
__device__ __forceinline__ int wrap1(int v, int size) {
    int w = v < 0 ? v + size : v;
    w = w >= size ? w - size : w;
    if (w < 0 || w >= size) w = wrap_any(v, size);
    return w;
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ void wrap3(const Planes& P, int q[3]) {
#pragma unroll
    for (int b = 0; b < 3; b++)
        if (P.per & (1u << b)) q[b] = wrap1(q[b], P.n[b]);
}
__device__ __forceinline__ bool valid_cell(const Planes& P, const int q[3]) {
    return q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < P.n[0] && q[1] < P.n[1] && q[2] < P.n[2];
}
__device__ __forceinline__ uint32_t cell_index(const Planes& P, const int q[3]) {
    return (uint32_t)q[0] + (uint32_t)P.n[0] * ((uint32_t)q[1] + (uint32_t)P.n[1] * (uint32_t)q[2]);
}
template <int A>
__device__ __forceinline__ uint32_t face_index(const Planes& P, const int q[3]) {
    const uint32_t px = (uint32_t)P.n[0] + (A == 0), py = (uint32_t)P.n[1] + (A == 1);
    const uint32_t off = A == 0 ? 0u : (A == 1 ? P.f0 : P.f0 + P.f1);
    return off + (uint32_t)q[0] + px * ((uint32_t)q[1] + py * (uint32_t)q[2]);
}
// open_face(q, A) (solver.wgsl:31) from the open bytes; q is wrapped in place.  A face at the
// upper end of a closed axis has no cell q and is closed.
template <int A>
__device__ __forceinline__ bool open_face(const Planes& P, const uint8_t* __restrict__ open8, int q[3]) {
    wrap3(P, q);
    const bool valid = valid_cell(P, q);
    // the byte is read from a clamped (always valid) cell so that no branch guards the load
    const int c[3] = {clampi(q[0], 0, P.n[0] - 1), clampi(q[1], 0, P.n[1] - 1), clampi(q[2], 0, P.n[2] - 1)};
    return valid && ((open8[cell_index(P, c)] >> A) & 1u);
}
// Faces of plane A, one wave per row along x (consecutive lanes, consecutive faces): each block
// takes a contiguous range of rows (y, z) so that stencil neighbours in y and z stay in its
// caches; f(x, y, z, k) with k the face index within the plane (x + px * row).
template <int A, class F>
__device__ __forceinline__ void for_faces(const Planes& P, const VGrid& vg, F&& f) {
    const uint32_t px = (uint32_t)P.n[0] + (A == 0), py = (uint32_t)P.n[1] + (A == 1);
    const uint32_t rows = py * ((uint32_t)P.n[2] + (A == 2));
    constexpr uint32_t kWaves = MHDP_BLOCK / 64;
    const uint32_t per = (rows + vg.n - 1) / vg.n;
    const uint32_t r0 = vg.b * per, r1 = min(rows, r0 + per);
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t row = r0 + (threadIdx.x >> 6); row < r1; row += kWaves) {
        const int y = (int)(row % py), z = (int)(row / py);
        for (uint32_t x = lane; x < px; x += 64) f((int)x, y, z, x + px * row);
    }
}
// floor()ed sample coordinate -> integer, clamped to +-2^30 first (NaN -> -2^30: fmaxf returns
// the number); such positions lie far outside any grid
__device__ __forceinline__ int cell_coord(float v) { return (int)fminf(fmaxf(v, -1073741824.0f), 1073741824.0f); }
// Trilinear samples of component C from the open-face copy, in two halves so that a face's
// independent samples have all their loads in flight together: corners_load works out the corner
// coordinates per axis (periodic: wrapped; closed: out-of-range corners read a clamped face and
// count as 0 - both formed and selected, no branch on the axis kind) and issues the eight loads;
// corners_sum adds the corners in the reference's order (z, y, x nested; value * wx * wy * wz).
struct Corners {
    float f[8];   // x fastest
    bool ok[3][2];
};
template <int C>
__device__ __forceinline__ void corners_load(const Planes& P, const float* __restrict__ uo, const int base[3],
                                             Corners& k) {
    const uint32_t px = (uint32_t)P.n[0] + (C == 0), py = (uint32_t)P.n[1] + (C == 1);
    const uint32_t stride[3] = {1u, px, px * py};
    uint32_t idx = C == 0 ? 0u : (C == 1 ? P.f0 : P.f0 + P.f1);  // index of the (0,0,0) corner
    uint32_t step[3];                                              // corner 1 minus corner 0
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const int n = P.n[b], lim = n + (b == C ? 1 : 0);
        int c0, c1;
        if ((P.per >> b) & 1u) {  // uniform over the launch: closed axes skip the wrap entirely
            c0 = wrap1(base[b], n);
            c1 = c0 + 1 == n ? 0 : c0 + 1;
            k.ok[b][0] = k.ok[b][1] = true;
        } else {
            c0 = clampi(base[b], 0, lim - 1);
            c1 = clampi(base[b] + 1, 0, lim - 1);
            k.ok[b][0] = base[b] >= 0 && base[b] < lim;
            k.ok[b][1] = base[b] + 1 >= 0 && base[b] + 1 < lim;
        }
        idx += (uint32_t)c0 * stride[b];
        step[b] = (uint32_t)(c1 - c0) * stride[b];
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
        k.f[j] = uo[idx + ((j & 1) ? step[0] : 0u) + ((j & 2) ? step[1] : 0u) + ((j & 4) ? step[2] : 0u)];
}
// pins the loaded values in registers: the loads stay issued ahead of any arithmetic
__device__ __forceinline__ void corners_fence(Corners& k) {
#pragma unroll
    for (int j = 0; j < 8; j++) asm volatile("" : "+v"(k.f[j]));
}
__device__ __forceinline__ float corners_sum(const Corners& k, const float frac[3]) {
    float value = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int x = j & 1, y = (j >> 1) & 1, z = j >> 2;
        const float wx = x ? frac[0] : 1.0f - frac[0];
        const float wy = y ? frac[1] : 1.0f - frac[1];
        const float wz = z ? frac[2] : 1.0f - frac[2];
        value += (k.ok[0][x] && k.ok[1][y] && k.ok[2][z] ? k.f[j] : 0.0f) * wx * wy * wz;
    }
    return value;
}
// sample_component (solver.wgsl:109-116): base cell and fractions of an arbitrary position
template <int C>
__device__ __forceinline__ void sample_coords(const float position[3], int base[3], float frac[3]) {
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const float q = position[b] - (b == C ? 0.0f : 0.5f);
        const float fl = floorf(q);
        base[b] = cell_coord(fl);
        frac[b] = q - fl;
    }
}
// ... and of the centre of face pos of plane A, where every step is exact and known: on axis b
// the sample coordinate is pos + 1/2 - 1/2 [b == A] - 1/2 [b != C], so the fraction is 1/2 when
// exactly one of b == A, b == C holds (else 0) and the base is pos - [b == A and b != C]
template <int C, int A>
__device__ __forceinline__ void centre_coords(const int pos[3], int base[3], float frac[3]) {
#pragma unroll
    for (int b = 0; b < 3; b++) {
        base[b] = pos[b] - ((b == A && b != C) ? 1 : 0);
        frac[b] = (b == A) != (b == C) ? 0.5f : 0.0f;
    }
}

// open bytes: bit a of cell i set when i and its lower neighbour on axis a (wrapped) are fluid;
// bit 3 set when i is fluid, bit 4 when i is its component's anchor (gauge mask 2 | root << 2)
__global__ void __launch_bounds__(MHDP_BLOCK) k_open8(Planes P, uint32_t n, const uint32_t* __restrict__ mask,
                                                      uint8_t* __restrict__ open8) {
    uint32_t lo, hi;
    block_range(n, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
        uint32_t bits = 0;
        if (mask[i]) {
            bits = (mask[i] & 3u) == 2u ? 24u : 8u;
            const uint32_t nxy = (uint32_t)P.n[0] * (uint32_t)P.n[1];
            int pos[3];
            pos[2] = (int)(i / nxy);
            const uint32_t r = i - (uint32_t)pos[2] * nxy;
            pos[1] = (int)(r / (uint32_t)P.n[0]);
            pos[0] = (int)(r - (uint32_t)pos[1] * (uint32_t)P.n[0]);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                int l[3] = {pos[0], pos[1], pos[2]};
                l[a] -= 1;
                wrap3(P, l);
                if (valid_cell(P, l) && mask[cell_index(P, l)] != 0u) bits |= 1u << a;
            }
        }
        open8[i] = (uint8_t)bits;
    }
}

template <int A>
struct K_open_faces {
    static constexpr int MINB = 0;
    struct Args {
        Planes P;
        const uint8_t* open8;
        const float* u;
        float* uo;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const Planes P = a.P;
        const uint32_t off = A == 0 ? 0u : (A == 1 ? P.f0 : P.f0 + P.f1);
        const uint8_t* __restrict__ open8 = a.open8;
        const float* __restrict__ u = a.u;
        float* __restrict__ uo = a.uo;
        for_faces<A>(P, vg, [&](int x, int y, int z, uint32_t k) {
            int q[3] = {x, y, z};
            float v = 0.0f;
            if (open_face<A>(P, open8, q)) v = u[face_index<A>(P, q)];
            uo[off + k] = v;
        });
    }
};

template <int A>
struct K_predict {
    static constexpr int MINB = 0;
    struct Args {
        Planes P;
        float s_half, s_full;  // 0.5 dt / h, dt / h
        const uint8_t* open8;
        const float* uo;
        float* out;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const Planes P = a.P;
        const uint32_t off = A == 0 ? 0u : (A == 1 ? P.f0 : P.f0 + P.f1);
        const uint8_t* __restrict__ open8 = a.open8;
        const float* __restrict__ uo = a.uo;
        float* __restrict__ out = a.out;
        for_faces<A>(P, vg, [&](int x, int y, int z, uint32_t k) {
            const int pos[3] = {x, y, z};
            int q[3] = {x, y, z};
            float result = 0.0f;
            if (open_face<A>(P, open8, q)) {
                {
                    int b0[3], b1[3], b2[3];
                    float r0[3], r1[3], r2[3], v[3], at[3], mid[3], dep[3];
                    Corners k0, k1, k2;
                    // velocity at the face centre (three components, 24 loads in flight)
                    centre_coords<0, A>(pos, b0, r0);
                    centre_coords<1, A>(pos, b1, r1);
                    centre_coords<2, A>(pos, b2, r2);
                    corners_load<0>(P, uo, b0, k0);
                    corners_load<1>(P, uo, b1, k1);
                    corners_load<2>(P, uo, b2, k2);
                    corners_fence(k0);
                    corners_fence(k1);
                    corners_fence(k2);
                    v[0] = corners_sum(k0, r0);
                    v[1] = corners_sum(k1, r1);
                    v[2] = corners_sum(k2, r2);
#pragma unroll
                    for (int b = 0; b < 3; b++) at[b] = (float)pos[b] + 0.5f;
                    at[A] -= 0.5f;
#pragma unroll
                    for (int b = 0; b < 3; b++) mid[b] = at[b] - v[b] * a.s_half;
                    // velocity at the half-step midpoint
                    sample_coords<0>(mid, b0, r0);
                    sample_coords<1>(mid, b1, r1);
                    sample_coords<2>(mid, b2, r2);
                    corners_load<0>(P, uo, b0, k0);
                    corners_load<1>(P, uo, b1, k1);
                    corners_load<2>(P, uo, b2, k2);
                    corners_fence(k0);
                    corners_fence(k1);
                    corners_fence(k2);
                    v[0] = corners_sum(k0, r0);
                    v[1] = corners_sum(k1, r1);
                    v[2] = corners_sum(k2, r2);
#pragma unroll
                    for (int b = 0; b < 3; b++) dep[b] = at[b] - v[b] * a.s_full;
                    // component A at the departure point
                    sample_coords<A>(dep, b0, r0);
                    corners_load<A>(P, uo, b0, k0);
                    corners_fence(k0);
                    result = corners_sum(k0, r0);
                }
            }
            out[off + k] = result;
        });
    }
};

// (liquid fraction, curvature) of every cell from its packet (reserved.z, reserved.w)
__global__ void __launch_bounds__(MHDP_BLOCK) k_cell_pair(uint32_t n, const float* __restrict__ cells,
                                                          float2* __restrict__ ak) {
    uint32_t lo, hi;
    block_range(n, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK)
        ak[i] = make_float2(cells[16u * (size_t)i + 14u], cells[16u * (size_t)i + 15u]);
}

template <int A>
struct K_viscous {
    static constexpr int MINB = 0;
    struct Args {
        Planes P;
        float dt, nu_dt_h2, h, rho_l, rho_air, sigma, acc;
        int two_phase;
        const uint8_t* open8;
        const float* pred;
        const float2* ak;    // two-phase: per cell (liquid fraction, curvature), the packets' reserved.z, .w
        const float* em;     // NULL or 16 floats per cell, force at 8..10
        float* out;
    };
    // the wrapped cell of q clamped into the grid (for loads that are used only when q is valid)
    static __device__ __forceinline__ uint32_t clamped_cell(const Planes& P, int q[3], bool& valid) {
        wrap3(P, q);
        valid = valid_cell(P, q);
        const int c[3] = {clampi(q[0], 0, P.n[0] - 1), clampi(q[1], 0, P.n[1] - 1), clampi(q[2], 0, P.n[2] - 1)};
        q[0] = c[0];
        q[1] = c[1];
        q[2] = c[2];
        return cell_index(P, c);
    }
    static __device__ void run(const Args& a, const VGrid& vg) {
        const Planes P = a.P;
        const uint32_t off = A == 0 ? 0u : (A == 1 ? P.f0 : P.f0 + P.f1);
        const uint8_t* __restrict__ open8 = a.open8;
        const float* __restrict__ pred = a.pred;
        const float2* __restrict__ ak = a.ak;
        const float* __restrict__ em = a.em;
        float* __restrict__ out = a.out;
        const int nj = a.two_phase ? 1 : 7;  // two-phase grids have no viscous term
        for_faces<A>(P, vg, [&](int x, int y, int z, uint32_t k) {
            // the face (j = 0) and its neighbours j = 1 + 2 d + (sign > 0); every load below is
            // issued from an always-valid index before any value is used
            bool valid[7];
            uint32_t o8[7];
            float pv[7];
#pragma unroll
            for (int j = 0; j < 7; j++) {
                if (j >= nj) break;
                int q[3] = {x, y, z};
                if (j > 0) q[(j - 1) >> 1] += ((j - 1) & 1) ? 1 : -1;
                const uint32_t c = clamped_cell(P, q, valid[j]);
                o8[j] = open8[c];
                pv[j] = pred[face_index<A>(P, q)];
            }
            int l[3] = {x, y, z}, r[3] = {x, y, z};
            l[A] -= 1;
            bool vl, vr;
            const uint32_t cl = clamped_cell(P, l, vl), cr = clamped_cell(P, r, vr);
            float eml = 0.0f, emr = 0.0f, al = 0.0f, ar = 0.0f, kl = 0.0f, kr = 0.0f;
            if (em) {
                eml = em[16u * (size_t)cl + 8u + (uint32_t)A];
                emr = em[16u * (size_t)cr + 8u + (uint32_t)A];
            }
            if (a.two_phase) {
                const float2 l2 = ak[cl], r2 = ak[cr];
                al = l2.x;
                kl = l2.y;
                ar = r2.x;
                kr = r2.y;
            }
#pragma unroll
            for (int j = 0; j < 7; j++) {
                if (j >= nj) break;
                asm volatile("" : "+v"(o8[j]), "+v"(pv[j]));
            }
            asm volatile("" : "+v"(eml), "+v"(emr), "+v"(al), "+v"(ar), "+v"(kl), "+v"(kr));
            float result = 0.0f;
            if (valid[0] && ((o8[0] >> A) & 1u)) {  // open_face (solver.wgsl:31)
                const float center = pv[0];
                float laplace = 0.0f;
                if (!a.two_phase) {
#pragma unroll
                    for (int j = 1; j < 7; j++) {
                        const bool open = valid[j] && ((o8[j] >> A) & 1u);
                        float nb = open ? pv[j] : 0.0f;
                        if (!open && ((j - 1) >> 1) != A) nb = -center;  // no-slip reflection
                        laplace += nb - center;
                    }
                }
                float rho = a.rho_l, cap = 0.0f, beta = 1.0f;
                if (a.two_phase) {
                    // density, face_density (solver.wgsl:41-42), capillary_force (solver.wgsl:52-57)
                    // and mobility (solver.wgsl:43-46) of an open face: both cells exist
                    const float dl = a.rho_air * (1.0f - al) + a.rho_l * al;
                    const float dr = a.rho_air * (1.0f - ar) + a.rho_l * ar;
                    rho = 0.5f * (dl + dr);
                    const float wl = al * (1.0f - al), wr = ar * (1.0f - ar);
                    const float kk = wl + wr > 1e-6f ? (wl * kl + wr * kr) / fmaxf(wl + wr, 1e-20f) : 0.5f * (kl + kr);
                    cap = a.sigma * kk * (ar - al) / a.h / rho;
                    beta = a.rho_l / rho;
                    beta *= rho / a.rho_l;
                }
                // em_force (solver.wgsl:108): zero without an EM field
                const float acc = a.acc + ((vl ? eml : 0.0f) + (vr ? emr : 0.0f)) / (2.0f * rho) + cap;
                result = (center + a.nu_dt_h2 * laplace + a.dt * acc) * beta;
            }
            out[off + k] = result;
        });
    }
};

// ---------------------------------------------------------------- two-phase viscous stress
// diffuse_surface_stress (src/fluid/mod.rs:1146-1234) with the stress operator of
// src/fluid/viscosity.wgsl: normal stresses at cells, shear stresses at edges (odd ghosts at closed
// walls), divided by h and the face density.  Per cell the dynamic viscosity and the density are
// computed once (K_stress_cells, the reference's stress_coefficients).  Each operator evaluation
// is two passes: K_stress_tensor writes every stress once per lattice position (the reference's
// stress_tensor, widened to both component orders of each shear stress), then the face kernels
// gather the divergence from it.  The FP32 operation order is the reference's throughout.
//
// Tensor layout: positions p of the padded lattice, 0 <= p_b < n_b + 1 on closed axes (the upper
// layer carries the shear stresses at the upper walls) and 0 <= p_b < n_b on periodic ones, x
// fastest; nine arrays of that length: the normal stresses of axes 0..2 (zero outside the grid),
// then the shear stress of edge (i, j) evaluated in component order (i, j) at slot
// 3 + 2 (i + j - 1) + [i > j].  The direct operator (explicit substeps, stress_acceleration with
// cached = false) reads slot (i, j) for component i.  The cached operator (RKC2 stages,
// tensor_shear at viscosity.wgsl:60-67) reads the ascending order (i < j) at in-grid positions,
// so the tensor pass writes that value into both slots there; outside the grid it evaluates the
// requested order, as the direct operator does.
struct StressCtx {
    Planes P;
    float h, nu_rho, mu_air, rho_l, rho_air;
    const uint32_t* mask;
    const uint8_t* open8;  // bits 0..2 open lower faces, bit 3 fluid cell
    const float* mu;       // per cell: dynamic viscosity (0 outside the fluid)
    const float* rho;      // per cell: density mix(rho_air, rho_l, occupancy)
    int t[3];              // padded lattice dims
    uint32_t tlen;         // t0 t1 t2
    float* tensor;         // nine arrays of tlen floats
};
__device__ __forceinline__ bool s_fluid(const StressCtx& c, int q[3]) {
    wrap3(c.P, q);
    return valid_cell(c.P, q) && c.mask[cell_index(c.P, q)] != 0u;
}
__device__ __forceinline__ float s_mu(const StressCtx& c, const int p[3]) {
    int q[3] = {p[0], p[1], p[2]};
    wrap3(c.P, q);
    return valid_cell(c.P, q) ? c.mu[cell_index(c.P, q)] : 0.0f;
}
__device__ __forceinline__ float s_rho(const StressCtx& c, const int p[3]) {
    int q[3] = {p[0], p[1], p[2]};
    wrap3(c.P, q);
    return valid_cell(c.P, q) ? c.rho[cell_index(c.P, q)] : c.rho_air * (1.0f - 0.0f) + c.rho_l * 0.0f;
}
__device__ __forceinline__ float s_face_density(const StressCtx& c, const int q[3], int i) {
    int l[3] = {q[0], q[1], q[2]};
    l[i] -= 1;
    return 0.5f * (s_rho(c, l) + s_rho(c, q));
}
__device__ __forceinline__ bool s_open(const StressCtx& c, int axis, const int p[3]) {
    int q[3] = {p[0], p[1], p[2]};
    wrap3(c.P, q);
    if (!valid_cell(c.P, q)) return false;
    return (c.open8[cell_index(c.P, q)] >> axis) & 1u;
}
// coefficient_edge_mu (viscosity.wgsl:16-22), for the stability bound
__device__ float s_edge_mu(const StressCtx& c, const int q[3], int i, int j) {
    float total = 0.0f, count = 0.0f;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            int at[3] = {q[0], q[1], q[2]};
            at[i] -= a;
            at[j] -= b;
            int w[3] = {at[0], at[1], at[2]};
            if (s_fluid(c, w)) {
                total += s_mu(c, at);
                count += 1.0f;
            }
        }
    return total / fmaxf(count, 1.0f);
}
// the wrapped cell of q clamped into the grid (loads from it are used only when q is valid)
__device__ __forceinline__ uint32_t s_cell(const Planes& P, int q[3], bool& valid) {
    wrap3(P, q);
    valid = valid_cell(P, q);
#pragma unroll
    for (int b = 0; b < 3; b++) q[b] = clampi(q[b], 0, P.n[b] - 1);
    return cell_index(P, q);
}
// index of lattice position p in the tensor arrays (periodic axes wrapped, closed ones clamped)
__device__ __forceinline__ uint32_t s_tpos(const StressCtx& c, const int p[3]) {
    int q[3];
#pragma unroll
    for (int b = 0; b < 3; b++) q[b] = (c.P.per >> b) & 1u ? wrap1(p[b], c.P.n[b]) : clampi(p[b], 0, c.t[b] - 1);
    return (uint32_t)q[0] + (uint32_t)c.t[0] * ((uint32_t)q[1] + (uint32_t)c.t[1] * (uint32_t)q[2]);
}
__host__ __device__ constexpr int shear_slot(int i, int j) { return 3 + 2 * (i + j - 1) + (i > j ? 1 : 0); }

// stress_coefficients: per cell viscosity and density
__global__ void __launch_bounds__(MHDP_BLOCK) k_stress_cells(uint32_t n, const uint32_t* __restrict__ mask,
                                                             const float* __restrict__ cells, float nu_rho, float mu_air,
                                                             float rho_l, float rho_air, float* __restrict__ mu,
                                                             float* __restrict__ rho) {
    uint32_t lo, hi;
    block_range(n, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
        const float a = cells[16u * (size_t)i + 14u];
        const float t = fminf(fmaxf(a, 0.0f), 1.0f);
        mu[i] = mask[i] != 0u ? mu_air * (1.0f - t) + nu_rho * t : 0.0f;  // dynamic_mu (viscosity.wgsl:3-6)
        rho[i] = rho_air * (1.0f - a) + rho_l * a;                          // density (solver.wgsl:41)
    }
}

// viscosity_bound (viscosity.wgsl:83-98): per block the largest open-face row sum
template <int A>
struct K_stress_bound {
    static constexpr int MINB = 0;
    struct Args {
        StressCtx c;
        float* part;  // one float per block (>= 0)
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const StressCtx& c = a.c;
        float bound = 0.0f;
        for_faces<A>(c.P, vg, [&](int x, int y, int z, uint32_t) {
            const int q[3] = {x, y, z};
            if (!s_open(c, A, q)) return;
            int left[3] = {x, y, z};
            left[A] -= 1;
            float sum = s_mu(c, q) + s_mu(c, left);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if (j == A) continue;
                int upper[3] = {x, y, z};
                upper[j] += 1;
                sum += s_edge_mu(c, q, A, j) + s_edge_mu(c, upper, A, j);
            }
            bound = fmaxf(bound, 4.0f * sum / (s_face_density(c, q, A) * c.h * c.h));
        });
        for (int o = 32; o > 0; o >>= 1) bound = fmaxf(bound, __shfl_down(bound, o, 64));
        __shared__ float red[MHDP_BLOCK / 64];
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = bound;
        __syncthreads();
        if (threadIdx.x == 0) {
            float b = red[0];
            for (int w = 1; w < MHDP_BLOCK / 64; w++) b = fmaxf(b, red[w]);
            a.part[vg.b] = b;
        }
    }
};
__global__ void k_stress_bound_fold(const float* __restrict__ part, uint32_t count, float* __restrict__ out) {
    __shared__ float red[64];
    float b = 0.0f;
    for (uint32_t i = threadIdx.x; i < count; i += 64) b = fmaxf(b, part[i]);
    red[threadIdx.x] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = red[0];
        for (int w = 1; w < 64; w++) m = fmaxf(m, red[w]);
        *out = m;
    }
}

// stress_tensor (viscosity.wgsl:30-55) on the padded lattice.  Per position the ten cells
// p, p - e_a, p - e_i - e_j and p + e_a supply the open/fluid bytes, the seven lower ones the
// viscosities, and twelve face values cover the three normal stresses (faces p and p + e_a on
// axis a) and the three edges (faces p and p - e_j on axis i); every load is issued from a
// clamped index before any value is used.
struct K_stress_tensor {
    static constexpr int MINB = 0;
    struct Args {
        StressCtx c;
        int cached;  // write the ascending-order shear into both slots at in-grid positions
        const float* u;
    };
    // cell k of the stencil: 0 p, 1 + a p - e_a, 4 + k p - e_i - e_j (pairs 01, 02, 12), 7 + a p + e_a
    static __device__ __forceinline__ void offset(int k, int q[3]) {
        if (k >= 1 && k <= 3) q[k - 1] -= 1;
        if (k >= 4 && k <= 6) {
            q[k == 6 ? 1 : 0] -= 1;
            q[k == 4 ? 1 : 2] -= 1;
        }
        if (k >= 7) q[k - 7] += 1;
    }
    static __device__ void run(const Args& a, const VGrid& vg) {
        const StressCtx& c = a.c;
        const Planes& P = c.P;
        const float* __restrict__ u = a.u;
        const uint32_t tx = (uint32_t)c.t[0], ty = (uint32_t)c.t[1];
        const uint32_t rows = ty * (uint32_t)c.t[2];
        constexpr uint32_t kWaves = MHDP_BLOCK / 64;
        const uint32_t per = (rows + vg.n - 1) / vg.n;
        const uint32_t r0 = vg.b * per, r1 = min(rows, r0 + per);
        for (uint32_t row = r0 + (threadIdx.x >> 6); row < r1; row += kWaves) {
            const int y = (int)(row % ty), z = (int)(row / ty);
            for (uint32_t x = threadIdx.x & 63; x < tx; x += 64) {
                bool valid[10];
                uint32_t bits[10];
                float mu[7];
                float f[12];  // 0..2 face a at p, 3..5 face a at p + e_a, 6 + 2 i + (j > other) face i at p - e_j
#pragma unroll
                for (int k = 0; k < 10; k++) {
                    int q[3] = {(int)x, y, z};
                    offset(k, q);
                    const uint32_t cell = s_cell(P, q, valid[k]);
                    bits[k] = c.open8[cell];
                    if (k < 7) mu[k] = c.mu[cell];
                    if (k == 0) {
                        f[0] = u[face_index<0>(P, q)];
                        f[1] = u[face_index<1>(P, q)];
                        f[2] = u[face_index<2>(P, q)];
                    }
                    if (k >= 1 && k <= 3) {  // p - e_j: the two other axes' faces
                        const int j = k - 1;
                        if (j != 0) f[6 + 0 + (j == 2)] = u[face_index<0>(P, q)];
                        if (j != 1) f[8 + (j == 2)] = u[face_index<1>(P, q)];
                        if (j != 2) f[10 + (j == 1)] = u[face_index<2>(P, q)];
                    }
                    if (k == 7) f[3] = u[face_index<0>(P, q)];
                    if (k == 8) f[4] = u[face_index<1>(P, q)];
                    if (k == 9) f[5] = u[face_index<2>(P, q)];
                }
#pragma unroll
                for (int k = 0; k < 10; k++) asm volatile("" : "+v"(bits[k]));
#pragma unroll
                for (int k = 0; k < 7; k++) asm volatile("" : "+v"(mu[k]));
#pragma unroll
                for (int k = 0; k < 12; k++) asm volatile("" : "+v"(f[k]));
                auto open = [&](int k, int axis) { return valid[k] && ((bits[k] >> axis) & 1u); };
                auto fluid = [&](int k) { return valid[k] && ((bits[k] >> 3) & 1u); };
                int p[3] = {(int)x, y, z};
                wrap3(P, p);
                const bool inside = valid_cell(P, p);
                const uint32_t t = (uint32_t)x + tx * ((uint32_t)y + ty * (uint32_t)z);
                float* __restrict__ out = c.tensor;
                // normal_stress (viscosity.wgsl:30-33)
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    const float lo = open(0, ax) ? f[ax] : 0.0f, hi = open(7 + ax, ax) ? f[3 + ax] : 0.0f;
                    out[(size_t)ax * c.tlen + t] = inside && fluid(0) ? 2.0f * mu[0] * (hi - lo) / c.h : 0.0f;
                }
                // shear_stress (viscosity.wgsl:34-45) of the edges (i, j), i < j, in both orders
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    const int i = e == 2 ? 1 : 0, j = e == 0 ? 1 : 2;
                    // face i at p and at p - e_j, face j at p and at p - e_i
                    const bool oq_i = open(0, i), ol_i = open(1 + j, i), oq_j = open(0, j), ol_j = open(1 + i, j);
                    float ui1 = oq_i ? f[i] : 0.0f, ui0 = ol_i ? f[6 + 2 * i + (j > (i == 0 ? 1 : 0))] : 0.0f;
                    float uj1 = oq_j ? f[j] : 0.0f, uj0 = ol_j ? f[6 + 2 * j + (i > (j == 0 ? 1 : 0))] : 0.0f;
                    if (!oq_i && ol_i) ui1 = -ui0;
                    if (!ol_i && oq_i) ui0 = -ui1;
                    if (!oq_j && ol_j) uj1 = -uj0;
                    if (!ol_j && oq_j) uj0 = -uj1;
                    // coefficient_edge_mu: cells p, p - e_j, p - e_i, p - e_i - e_j for order (i, j)
                    const int cj = 1 + j, ci = 1 + i, cij = 4 + e;
                    float tij = 0.0f, tji = 0.0f, count = 0.0f;
                    if (fluid(0)) { tij += mu[0]; tji += mu[0]; count += 1.0f; }
                    if (fluid(cj)) { tij += mu[cj]; count += 1.0f; }
                    if (fluid(ci)) { tij += mu[ci]; tji += mu[ci]; count += 1.0f; }
                    if (fluid(cj)) tji += mu[cj];
                    if (fluid(cij)) { tij += mu[cij]; tji += mu[cij]; count += 1.0f; }
                    const float sij = tij / fmaxf(count, 1.0f) * (ui1 - ui0 + uj1 - uj0) / c.h;
                    const float sji = tji / fmaxf(count, 1.0f) * (uj1 - uj0 + ui1 - ui0) / c.h;
                    out[(size_t)shear_slot(i, j) * c.tlen + t] = sij;
                    out[(size_t)shear_slot(j, i) * c.tlen + t] = a.cached && inside ? sij : sji;
                }
            }
        }
    }
};

// the operator value of face q of plane A from the tensor: stress_acceleration
// (viscosity.wgsl:68-82), loads first (StressFace::load), arithmetic after (acc)
template <int A>
struct StressFace {
    static constexpr int J0 = A == 0 ? 1 : 0, J1 = A == 2 ? 1 : 2;
    bool open, valid_l;
    uint32_t canon;  // face index of the wrapped face
    float u, nq, nl, s0q, s0u, s1q, s1u, rq, rl;
    __device__ __forceinline__ void load(const StressCtx& c, const float* __restrict__ uin, int x, int y, int z) {
        const float* __restrict__ T = c.tensor;
        int q[3] = {x, y, z}, l[3] = {x, y, z}, u0[3] = {x, y, z}, u1[3] = {x, y, z};
        l[A] -= 1;
        u0[J0] += 1;
        u1[J1] += 1;
        s0u = T[(size_t)shear_slot(A, J0) * c.tlen + s_tpos(c, u0)];
        s1u = T[(size_t)shear_slot(A, J1) * c.tlen + s_tpos(c, u1)];
        const uint32_t tq = s_tpos(c, q), tl = s_tpos(c, l);
        nq = T[(size_t)A * c.tlen + tq];
        nl = T[(size_t)A * c.tlen + tl];
        s0q = T[(size_t)shear_slot(A, J0) * c.tlen + tq];
        s1q = T[(size_t)shear_slot(A, J1) * c.tlen + tq];
        bool valid_q;
        const uint32_t cq = s_cell(c.P, q, valid_q), cl = s_cell(c.P, l, valid_l);
        canon = face_index<A>(c.P, q);
        const uint32_t bits = c.open8[cq];
        u = uin[canon];
        rq = c.rho[cq];
        rl = c.rho[cl];
        open = valid_q && ((bits >> A) & 1u);
    }
    __device__ __forceinline__ void fence() {
        asm volatile("" : "+v"(u), "+v"(nq), "+v"(nl), "+v"(s0q), "+v"(s0u), "+v"(s1q), "+v"(s1u));
        asm volatile("" : "+v"(rq), "+v"(rl));
    }
    // open faces only: both cells of the face exist (the face density is then the two densities)
    __device__ __forceinline__ float acc(const StressCtx& c) const {
        float div = nq - nl;
        div += s0u - s0q;
        div += s1u - s1q;
        return div / (c.h * (0.5f * (rl + rq)));
    }
};

// stress_diffuse (viscosity.wgsl:110-117): one explicit substep (tensor of u, direct form).  A face
// reads u only at its own canonical index (the rest comes from the tensor), and on a periodic axis
// the thread of face 0 also writes its duplicate at n_A (which nobody reads), so out may be u.
template <int A>
struct K_stress_euler {
    static constexpr int MINB = 0;
    struct Args {
        StressCtx c;
        float dt;
        const float* u;
        float* out;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const StressCtx& c = a.c;
        const uint32_t off = A == 0 ? 0u : (A == 1 ? c.P.f0 : c.P.f0 + c.P.f1);
        const bool per = (c.P.per >> A) & 1u;
        const uint32_t px = (uint32_t)c.P.n[0] + (A == 0), py = (uint32_t)c.P.n[1] + (A == 1);
        const uint32_t dup = (uint32_t)c.P.n[A] * (A == 0 ? 1u : (A == 1 ? px : px * py));
        for_faces<A>(c.P, vg, [&](int x, int y, int z, uint32_t k) {
            const int qa = A == 0 ? x : (A == 1 ? y : z);
            if (per && qa == c.P.n[A]) return;  // written with face 0
            StressFace<A> s;
            s.load(c, a.u, x, y, z);
            s.fence();
            const float v = s.open ? s.u + a.dt * s.acc(c) : 0.0f;
            a.out[off + k] = v;
            if (per && qa == 0) a.out[off + k + dup] = v;
        });
    }
};

// stress_save_initial's operator value (stabilized.wgsl:7-16), per face of the plane (cached tensor)
template <int A>
struct K_stress_initial {
    static constexpr int MINB = 0;
    struct Args {
        StressCtx c;
        const float* u;
        float* f0;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const StressCtx& c = a.c;
        const uint32_t off = A == 0 ? 0u : (A == 1 ? c.P.f0 : c.P.f0 + c.P.f1);
        for_faces<A>(c.P, vg, [&](int x, int y, int z, uint32_t k) {
            StressFace<A> s;
            s.load(c, a.u, x, y, z);
            s.fence();
            a.f0[off + k] = s.open ? s.acc(c) : 0.0f;
        });
    }
};

// stress_stabilized (stabilized.wgsl:16-30): one RKC2 stage (cached tensor of prev); out may be
// older (each face reads only its own element of older)
template <int A>
struct K_stress_stage {
    static constexpr int MINB = 0;
    struct Args {
        StressCtx c;
        float r0, r1, r2, r3, r4;
        const float* prev;
        const float* older;
        const float* y0;
        const float* f0;
        float* out;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const StressCtx& c = a.c;
        const uint32_t off = A == 0 ? 0u : (A == 1 ? c.P.f0 : c.P.f0 + c.P.f1);
        for_faces<A>(c.P, vg, [&](int x, int y, int z, uint32_t k) {
            StressFace<A> s;
            s.load(c, a.prev, x, y, z);
            float older = a.older[off + k], y0 = a.y0[s.canon], f0 = a.f0[s.canon];
            s.fence();
            asm volatile("" : "+v"(older), "+v"(y0), "+v"(f0));
            float v = 0.0f;
            if (s.open) v = a.r0 * s.u + a.r1 * older + a.r2 * y0 + a.r3 * s.acc(c) + a.r4 * f0;
            a.out[off + k] = v;
        });
    }
};

// Block count of the step kernels (measured at 512^3): enough resident waves to overlap the
// dependent load chains (open byte -> face value -> samples), fewer where a kernel reads the
// 64-byte cell packets and a block's rows must share the cache; MHDP_STEP_CAP overrides it.
uint32_t step_blocks(uint32_t faces, uint32_t cap) {
    static const long env = [] {
        const char* e = std::getenv("MHDP_STEP_CAP");
        return e ? std::strtol(e, nullptr, 10) : 0L;
    }();
    if (env >= 8 && env <= 65536) cap = (uint32_t)env;
    return grid_blocks(faces, cap);
}

}  // namespace

void launch_open8(hipStream_t s, const Grid3& g, const uint32_t* mask, uint8_t* open8) {
    hipLaunchKernelGGL(k_open8, dim3(grid_blocks(g.n)), dim3(MHDP_BLOCK), 0, s, planes_of(g), g.n, mask, open8);
}

void launch_predict(hipStream_t s, const Grid3& g, float h, float dt, int advect, const uint8_t* open8,
                    const float* u, float* uo, float* out) {
    const Planes P = planes_of(g);
    // without advection the result is the open-face copy itself (it may overwrite u: a face and
    // the periodic duplicate that reads it have the same open state and value)
    float* copy = advect ? uo : out;
    go<K_open_faces<0>>(s, step_blocks(plane_size(P, 0), 2048), {P, open8, u, copy});
    go<K_open_faces<1>>(s, step_blocks(plane_size(P, 1), 2048), {P, open8, u, copy});
    go<K_open_faces<2>>(s, step_blocks(plane_size(P, 2), 2048), {P, open8, u, copy});
    if (!advect) return;
    const float s_half = 0.5f * dt / h, s_full = dt / h;
    go<K_predict<0>>(s, step_blocks(plane_size(P, 0), 8192), {P, s_half, s_full, open8, uo, out});
    go<K_predict<1>>(s, step_blocks(plane_size(P, 1), 8192), {P, s_half, s_full, open8, uo, out});
    go<K_predict<2>>(s, step_blocks(plane_size(P, 2), 8192), {P, s_half, s_full, open8, uo, out});
}

void launch_viscous(hipStream_t s, const Grid3& g, float h, float rho_l, float rho_air, int two_phase, float dt,
                    float viscosity, const float acc[3], float sigma, const uint8_t* open8, const float* pred,
                    const float* cells, float2* ak, const float* em, float* out) {
    const Planes P = planes_of(g);
    const float k = viscosity * dt / (h * h);
    // two-phase: the two packet words each face needs of both its cells, gathered once (the three
    // plane kernels then read 8 bytes per cell instead of a 64-byte packet line each)
    if (two_phase) hipLaunchKernelGGL(k_cell_pair, dim3(grid_blocks(g.n)), dim3(MHDP_BLOCK), 0, s, g.n, cells, ak);
    const uint32_t cap = em ? 1024 : 8192;
    go<K_viscous<0>>(s, step_blocks(plane_size(P, 0), cap),
                     {P, dt, k, h, rho_l, rho_air, sigma, acc[0], two_phase, open8, pred, ak, em, out});
    go<K_viscous<1>>(s, step_blocks(plane_size(P, 1), cap),
                     {P, dt, k, h, rho_l, rho_air, sigma, acc[1], two_phase, open8, pred, ak, em, out});
    go<K_viscous<2>>(s, step_blocks(plane_size(P, 2), cap),
                     {P, dt, k, h, rho_l, rho_air, sigma, acc[2], two_phase, open8, pred, ak, em, out});
}

namespace {
StressCtx stress_ctx(const Grid3& g, const StressBufs& b) {
    StressCtx c;
    c.P = planes_of(g);
    c.h = b.h;
    c.nu_rho = b.nu * b.rho_l;
    c.mu_air = b.mu_air;
    c.rho_l = b.rho_l;
    c.rho_air = b.rho_air;
    c.mask = b.mask;
    c.open8 = b.open8;
    c.mu = b.mu;
    c.rho = b.rho;
    for (int a = 0; a < 3; a++) c.t[a] = c.P.n[a] + ((g.per >> a) & 1u ? 0 : 1);
    c.tlen = (uint32_t)stress_tensor_len(g) / 9u;
    c.tensor = b.tensor;
    return c;
}
}  // namespace

size_t stress_tensor_len(const Grid3& g) {
    size_t len = 9;
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; a++) len *= (size_t)n[a] + ((g.per >> a) & 1u ? 0 : 1);
    return len;
}

void launch_stress_cells(hipStream_t s, const Grid3& g, const uint32_t* mask, const float* cells, float nu, float mu_air,
                         float rho_l, float rho_air, float* mu, float* rho) {
    hipLaunchKernelGGL(k_stress_cells, dim3(grid_blocks(g.n)), dim3(MHDP_BLOCK), 0, s, g.n, mask, cells, nu * rho_l,
                       mu_air, rho_l, rho_air, mu, rho);
}

uint32_t stress_bound_blocks(const Grid3& g) {
    const Planes P = planes_of(g);
    return step_blocks(plane_size(P, 0), 1024) + step_blocks(plane_size(P, 1), 1024) + step_blocks(plane_size(P, 2), 1024);
}

void launch_stress_bound(hipStream_t s, const Grid3& g, const StressBufs& b, float* part, float* out) {
    const StressCtx c = stress_ctx(g, b);
    const uint32_t b0 = step_blocks(plane_size(c.P, 0), 1024), b1 = step_blocks(plane_size(c.P, 1), 1024),
                   b2 = step_blocks(plane_size(c.P, 2), 1024);
    go<K_stress_bound<0>>(s, b0, {c, part});
    go<K_stress_bound<1>>(s, b1, {c, part + b0});
    go<K_stress_bound<2>>(s, b2, {c, part + b0 + b1});
    hipLaunchKernelGGL(k_stress_bound_fold, dim3(1), dim3(64), 0, s, part, b0 + b1 + b2, out);
}

void launch_stress_tensor(hipStream_t s, const Grid3& g, const StressBufs& b, int cached, const float* u) {
    const StressCtx c = stress_ctx(g, b);
    go<K_stress_tensor>(s, step_blocks(c.tlen, 4096), {c, cached, u});
}

void launch_stress_euler(hipStream_t s, const Grid3& g, const StressBufs& b, float dt, const float* u, float* out) {
    const StressCtx c = stress_ctx(g, b);
    go<K_stress_euler<0>>(s, step_blocks(plane_size(c.P, 0), 4096), {c, dt, u, out});
    go<K_stress_euler<1>>(s, step_blocks(plane_size(c.P, 1), 4096), {c, dt, u, out});
    go<K_stress_euler<2>>(s, step_blocks(plane_size(c.P, 2), 4096), {c, dt, u, out});
}

void launch_stress_initial(hipStream_t s, const Grid3& g, const StressBufs& b, const float* u, float* f0) {
    const StressCtx c = stress_ctx(g, b);
    go<K_stress_initial<0>>(s, step_blocks(plane_size(c.P, 0), 4096), {c, u, f0});
    go<K_stress_initial<1>>(s, step_blocks(plane_size(c.P, 1), 4096), {c, u, f0});
    go<K_stress_initial<2>>(s, step_blocks(plane_size(c.P, 2), 4096), {c, u, f0});
}

void launch_stress_stage(hipStream_t s, const Grid3& g, const StressBufs& b, const float row[5], const float* prev,
                         const float* older, const float* y0, const float* f0, float* out) {
    const StressCtx c = stress_ctx(g, b);
    go<K_stress_stage<0>>(s, step_blocks(plane_size(c.P, 0), 4096),
                          {c, row[0], row[1], row[2], row[3], row[4], prev, older, y0, f0, out});
    go<K_stress_stage<1>>(s, step_blocks(plane_size(c.P, 1), 4096),
                          {c, row[0], row[1], row[2], row[3], row[4], prev, older, y0, f0, out});
    go<K_stress_stage<2>>(s, step_blocks(plane_size(c.P, 2), 4096),
                          {c, row[0], row[1], row[2], row[3], row[4], prev, older, y0, f0, out});
}
