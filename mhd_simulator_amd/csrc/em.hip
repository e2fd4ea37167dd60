// The EM step's Jacobi-CG on the edge curl-curl operator (src/em/induction.wgsl; the loop of
// EmState::advance_to, src/em/mod.rs:863-1010) with its increment prediction and the nodal
// auxiliary-space correction it switches to on long induction solves (src/em/nodal.wgsl).
// Built with -ffp-contract=off: the operator, right-hand side and coefficients keep the
// reference's FP32 operation order and equal the oracle's bit for bit.  Reductions are FP64
// per-block partials folded in block order by one block (deterministic).
//
// Edges of axis a have dims n + 1 except n_a on axis a, packed x | y | z; faces of axis a have
// dims n + e_a.  Per edge: sigma (mean conductivity of the four cells), diag (the four face
// reluctivities, 1 on boundary edges, + 1 per air endpoint), flags (1 boundary, 2 lower air node,
// 4 upper air node, 8 unit-reluctivity air stencil); per face: nu (kept instead of recomputing it
// from the material table at every product).
#include "launch.h"
#include "mhdp_kernels.h"

namespace {

enum : uint32_t { kBoundary = 1, kLowerAir = 2, kUpperAir = 4, kScalarAir = 8 };

struct EmGeo {
    int n[3];
    uint32_t E[3], F[3];
    uint32_t Etot;
    float h;
};

__device__ __forceinline__ void edge_dims(const EmGeo& g, int axis, uint32_t d[3]) {
#pragma unroll
    for (int b = 0; b < 3; b++) d[b] = (uint32_t)g.n[b] + (b == axis ? 0u : 1u);
}
__device__ __forceinline__ bool edge_valid(const EmGeo& g, int axis, const int p[3]) {
    uint32_t d[3];
    edge_dims(g, axis, d);
    return p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && (uint32_t)p[0] < d[0] && (uint32_t)p[1] < d[1] &&
           (uint32_t)p[2] < d[2];
}
__device__ __forceinline__ uint32_t edge_index(const EmGeo& g, int axis, const int p[3]) {
    uint32_t d[3];
    edge_dims(g, axis, d);
    const uint32_t off = axis == 0 ? 0u : (axis == 1 ? g.E[0] : g.E[0] + g.E[1]);
    return off + (uint32_t)p[0] + d[0] * ((uint32_t)p[1] + d[1] * (uint32_t)p[2]);
}
__device__ __forceinline__ int edge_decode(const EmGeo& g, uint32_t i, int p[3]) {
    int axis = 0;
    if (i >= g.E[0]) {
        i -= g.E[0];
        axis = 1;
        if (i >= g.E[1]) {
            i -= g.E[1];
            axis = 2;
        }
    }
    uint32_t d[3];
    edge_dims(g, axis, d);
    p[0] = (int)(i % d[0]);
    p[1] = (int)((i / d[0]) % d[1]);
    p[2] = (int)(i / (d[0] * d[1]));
    return axis;
}
__device__ __forceinline__ bool face_valid(const EmGeo& g, int axis, const int p[3]) {
    return p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] < g.n[0] + (axis == 0) && p[1] < g.n[1] + (axis == 1) &&
           p[2] < g.n[2] + (axis == 2);
}
__device__ __forceinline__ uint32_t face_index(const EmGeo& g, int axis, const int p[3]) {
    const uint32_t off = axis == 0 ? 0u : (axis == 1 ? g.F[0] : g.F[0] + g.F[1]);
    const uint32_t px = (uint32_t)g.n[0] + (axis == 0), py = (uint32_t)g.n[1] + (axis == 1);
    return off + (uint32_t)p[0] + px * ((uint32_t)p[1] + py * (uint32_t)p[2]);
}
__device__ __forceinline__ bool boundary(const EmGeo& g, int axis, const int p[3]) {
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    return p[a] == 0 || p[a] == g.n[a] || p[b] == 0 || p[b] == g.n[b];
}

// a_edge (induction.wgsl:33-38)
__device__ __forceinline__ float edge_val(const EmGeo& g, const float* __restrict__ v, int axis, const int p[3]) {
    return edge_valid(g, axis, p) ? v[edge_index(g, axis, p)] : 0.0f;
}
// curl_face (induction.wgsl:40-44)
__device__ __forceinline__ float curl(const EmGeo& g, const float* __restrict__ v, int axis, const int p[3]) {
    if (!face_valid(g, axis, p)) return 0.0f;
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    int qa[3] = {p[0], p[1], p[2]}, qb[3] = {p[0], p[1], p[2]};
    qa[a] += 1;
    qb[b] += 1;
    return edge_val(g, v, b, qa) - edge_val(g, v, b, p) - edge_val(g, v, a, qb) + edge_val(g, v, a, p);
}
// face_term (induction.wgsl:45-49)
__device__ __forceinline__ float face_term(const EmGeo& g, const float* __restrict__ nu, const float* __restrict__ v,
                                           int axis, const int p[3], bool contrast) {
    if (!face_valid(g, axis, p)) return 0.0f;
    float n = nu[face_index(g, axis, p)];
    if (contrast) n = 1.0f - n;
    return n * curl(g, v, axis, p);
}
// adjoint (induction.wgsl:51-55)
__device__ __forceinline__ float adjoint(const EmGeo& g, const float* __restrict__ nu, const float* __restrict__ v,
                                         int axis, const int p[3], bool contrast) {
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    int qa[3] = {p[0], p[1], p[2]}, qb[3] = {p[0], p[1], p[2]};
    qa[a] -= 1;
    qb[b] -= 1;
    return face_term(g, nu, v, b, p, contrast) - face_term(g, nu, v, b, qa, contrast) -
           face_term(g, nu, v, a, p, contrast) + face_term(g, nu, v, a, qb, contrast);
}
// node_div, gauge (induction.wgsl:64-71)
__device__ __forceinline__ float node_div(const EmGeo& g, const float* __restrict__ v, const int p[3]) {
    float value = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int q[3] = {p[0], p[1], p[2]};
        q[a] -= 1;
        value += edge_val(g, v, a, p) - edge_val(g, v, a, q);
    }
    return value;
}
__device__ __forceinline__ float gauge(const EmGeo& g, uint32_t fl, const float* __restrict__ v, int axis,
                                       const int p[3]) {
    int upper[3] = {p[0], p[1], p[2]};
    upper[axis] += 1;
    float value = 0.0f;
    if (fl & kLowerAir) value += node_div(g, v, p);
    if (fl & kUpperAir) value -= node_div(g, v, upper);
    return value;
}
using Phys = EmPhys;
// mass, diagonal (induction.wgsl:72-73)
__device__ __forceinline__ float mass(const EmGeo& g, const Phys& ph, float sigma) {
    if (ph.magnetostatic) return 0.0f;
    return ph.mu0 * sigma * g.h * g.h / ph.dt;
}
__device__ __forceinline__ float diagonal(const EmGeo& g, const Phys& ph, float diag, float sigma) {
    const float d = diag + mass(g, ph, sigma);
    return d > 1e-30f ? d : 1e-30f;
}
// air_laplacian (induction.wgsl:150-158)
__device__ __forceinline__ float air_laplacian(const EmGeo& g, const float* __restrict__ v, uint32_t i, int axis) {
    uint32_t d[3];
    edge_dims(g, axis, d);
    const uint32_t zs = d[0] * d[1];
    const float center = v[i];
    float result = 0.0f;
    result += (center - v[i - 1u]) + (center - v[i + 1u]);
    result += (center - v[i - d[0]]) + (center - v[i + d[0]]);
    result += (center - v[i - zs]) + (center - v[i + zs]);
    return result;
}

// cell_properties (induction.wgsl:16-25): (sigma, mu_r); props[2k], props[2k+1] for material k
__device__ __forceinline__ float2 cell_props(const EmGeo& g, const uint32_t* __restrict__ materials,
                                             const float* __restrict__ props, int vacuum, const int p[3]) {
    if (vacuum || p[0] < 0 || p[1] < 0 || p[2] < 0 || p[0] >= g.n[0] || p[1] >= g.n[1] || p[2] >= g.n[2])
        return make_float2(0.0f, 1.0f);
    const uint32_t id = materials[(uint32_t)p[0] + (uint32_t)g.n[0] * ((uint32_t)p[1] + (uint32_t)g.n[1] * (uint32_t)p[2])];
    if (id == 0u) return make_float2(props[0], props[1]);
    if (id & 0x80000000u) {
        const float alpha = (float)((id >> 8) & 0x7fffffu) / 8388607.0f;
        const uint32_t k = id & 255u;
        return make_float2(alpha * props[2 * k], props[1] * (1.0f - alpha) + props[2 * k + 1] * alpha);
    }
    return make_float2(props[2 * id], props[2 * id + 1]);
}
// nu_face (induction.wgsl:26-30)
__device__ __forceinline__ float nu_face(const EmGeo& g, const uint32_t* __restrict__ materials,
                                         const float* __restrict__ props, int vacuum, int axis, const int p[3]) {
    if (!face_valid(g, axis, p)) return 0.0f;
    int q[3] = {p[0], p[1], p[2]};
    q[axis] -= 1;
    return 0.5f / cell_props(g, materials, props, vacuum, p).y + 0.5f / cell_props(g, materials, props, vacuum, q).y;
}
// air_node (induction.wgsl:58-63)
__device__ __forceinline__ bool air_node(const EmGeo& g, const uint32_t* __restrict__ materials,
                                         const float* __restrict__ props, int vacuum, const int p[3]) {
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (p[a] <= 0 || p[a] >= g.n[a]) return false;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int z = 0; z < 2; z++) {
                const int q[3] = {p[0] - x, p[1] - y, p[2] - z};
                if (cell_props(g, materials, props, vacuum, q).x > 0.0f) return false;
            }
    return true;
}

// one block's four FP64 partials, in a fixed order (wave shuffles, then the waves)
// KEEP: how many components the block writes (the nodal passes replace only r.z and r.r of the
// records the pass before them wrote)
template <int OP1, int OP2, int KEEP = 4>  // 0 sum, 1 min, 2 max for components 1 and 2 (0 and 3 always sum)
__device__ void block_partials(double v[4], double* __restrict__ part, uint32_t b) {
    __shared__ double red[MHDP_BLOCK / 64][4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int op = k == 1 ? OP1 : (k == 2 ? OP2 : 0);
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) {
            const double w = __shfl_down(x, o, 64);
            x = op == 0 ? x + w : (op == 1 ? fmin(x, w) : fmax(x, w));
        }
        if (lane == 0) red[wid][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < KEEP) {
        const int k = threadIdx.x;
        const int op = k == 1 ? OP1 : (k == 2 ? OP2 : 0);
        double x = red[0][k];
        for (int w = 1; w < MHDP_BLOCK / 64; w++) x = op == 0 ? x + red[w][k] : (op == 1 ? fmin(x, red[w][k]) : fmax(x, red[w][k]));
        part[4u * b + (uint32_t)k] = x;
    }
}

struct K_em_nu {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        const uint32_t* materials;
        const float* props;
        int vacuum;
        float* nu;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const uint32_t total = a.g.F[0] + a.g.F[1] + a.g.F[2];
        uint32_t lo, hi;
        block_range(vg, total, lo, hi);
        for (uint32_t k = lo + threadIdx.x; k < hi; k += MHDP_BLOCK) {
            uint32_t j = k;
            int axis = 0;
            if (j >= a.g.F[0]) {
                j -= a.g.F[0];
                axis = 1;
                if (j >= a.g.F[1]) {
                    j -= a.g.F[1];
                    axis = 2;
                }
            }
            const uint32_t px = (uint32_t)a.g.n[0] + (axis == 0), py = (uint32_t)a.g.n[1] + (axis == 1);
            const int p[3] = {(int)(j % px), (int)((j / px) % py), (int)(j / (px * py))};
            a.nu[k] = nu_face(a.g, a.materials, a.props, a.vacuum, axis, p);
        }
    }
};

// coefficients (induction.wgsl:82-99); partials: boundary conductors, min sigma > 0, max sigma,
// conducting edges (coefficient_check)
struct K_em_coef {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        const uint32_t* materials;
        const float* props;
        int vacuum;
        float* sigma;
        float* diag;
        uint8_t* flags;
        double* part;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const EmGeo g = a.g;
        uint32_t lo, hi;
        block_range(vg, g.Etot, lo, hi);
        double v[4] = {0.0, 1e30, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int p[3];
            const int axis = edge_decode(g, i, p);
            const int ea = (axis + 1) % 3, eb = (axis + 2) % 3;
            float s = 0.0f;
            for (int u = 0; u < 2; u++)
                for (int w = 0; w < 2; w++) {
                    int q[3] = {p[0], p[1], p[2]};
                    q[ea] -= u;
                    q[eb] -= w;
                    s += 0.25f * cell_props(g, a.materials, a.props, a.vacuum, q).x;
                }
            int qa[3] = {p[0], p[1], p[2]}, qb[3] = {p[0], p[1], p[2]};
            qa[ea] -= 1;
            qb[eb] -= 1;
            const float n1 = nu_face(g, a.materials, a.props, a.vacuum, eb, p);
            const float n2 = nu_face(g, a.materials, a.props, a.vacuum, eb, qa);
            const float n3 = nu_face(g, a.materials, a.props, a.vacuum, ea, p);
            const float n4 = nu_face(g, a.materials, a.props, a.vacuum, ea, qb);
            float d = n1 + n2 + n3 + n4;
            uint32_t fl = 0;
            if (boundary(g, axis, p)) {
                d = 1.0f;
                fl |= kBoundary;
                if (s > 0.0f) v[0] += 1.0;
            }
            int upper[3] = {p[0], p[1], p[2]};
            upper[axis] += 1;
            const bool la = air_node(g, a.materials, a.props, a.vacuum, p);
            const bool ua = air_node(g, a.materials, a.props, a.vacuum, upper);
            if (la) fl |= kLowerAir;
            if (ua) fl |= kUpperAir;
            if (la && ua && n1 == 1.0f && n2 == 1.0f && n3 == 1.0f && n4 == 1.0f) fl |= kScalarAir;
            a.sigma[i] = s;
            a.diag[i] = d + (la ? 1.0f : 0.0f) + (ua ? 1.0f : 0.0f);
            a.flags[i] = (uint8_t)fl;
            if (s > 0.0f) {
                v[1] = fmin(v[1], (double)s);
                v[3] += 1.0;
            }
            v[2] = fmax(v[2], (double)s);
        }
        block_partials<1, 2>(v, a.part, vg.b);
    }
};

using Vecs = EmVecs;

// the step's right-hand side at edge i (prepare, induction.wgsl:121-124; with the increment x
// as well: true_residual, induction.wgsl:236-238)
__device__ __forceinline__ float em_rhs(const EmGeo& g, const Phys& ph, const Vecs& v, uint32_t i, int axis,
                                        const int p[3], bool with_x) {
    const uint32_t fl = v.flags[i];
    const float m = mass(g, ph, v.sigma[i]);
    const float s2 = v.mot ? v.mot[i] : 0.0f;
    const float held = with_x ? v.x[i] + (v.as_now[i] - v.as_prev[i]) : v.as_now[i] - v.as_prev[i];
    float b = m * ph.dt * s2 - m * held + adjoint(g, v.nu, v.as_now, axis, p, true) -
              adjoint(g, v.nu, v.a_prev, axis, p, false) - gauge(g, fl, v.a_prev, axis, p);
    if (with_x) b = b - adjoint(g, v.nu, v.x, axis, p, false) - gauge(g, fl, v.x, axis, p);
    if (fl & kBoundary) b = 0.0f;
    return b;
}

// prepare (induction.wgsl:116-129): from a zero increment (guess 0: x is cleared) or from the
// increment x already holds (guess 1: r = b - (C^T nu C x + gauge(x) + mass x), FP32 in that
// order); partials (rz, rr, bb, |b| before scaling)
struct K_em_prepare {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        Vecs v;
        EmScal* sc;
        double* part;
        int guess;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const EmGeo g = a.g;
        const float scale = (float)a.sc->scale;
        uint32_t lo, hi;
        block_range(vg, g.Etot, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int p[3];
            const int axis = edge_decode(g, i, p);
            float b = em_rhs(g, a.ph, a.v, i, axis, p, false);
            float r = b;
            if (a.guess) {
                const uint32_t fl = a.v.flags[i];
                if (!(fl & kBoundary)) {
                    const float ax = adjoint(g, a.v.nu, a.v.x, axis, p, false) + gauge(g, fl, a.v.x, axis, p);
                    r = b - (ax + mass(g, a.ph, a.v.sigma[i]) * a.v.x[i]);
                }
            } else {
                a.v.x[i] = 0.0f;
            }
            acc[3] += fabs((double)b);
            b = b / scale;
            r = r / scale;
            const float z = r / diagonal(g, a.ph, a.v.diag[i], a.v.sigma[i]);
            a.v.r[i] = r;
            a.v.p[i] = z;
            acc[0] += (double)r * (double)z;
            acc[1] += (double)r * (double)r;
            acc[2] += (double)b * (double)b;
        }
        block_partials<0, 0>(acc, a.part, vg.b);
    }
};

// begin_step (induction.wgsl:251-256): the increment's start prediction * (a_prev - a_before) on
// every non-boundary edge, 0 on the boundary
struct K_em_begin {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const uint8_t* flags;
        const float* a_prev;
        const float* a_before;
        float prediction;
        float* x;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        uint32_t lo, hi;
        block_range(vg, a.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK)
            a.x[i] = (a.flags[i] & kBoundary) ? 0.0f : a.prediction * (a.a_prev[i] - a.a_before[i]);
    }
};

// The operator product in two passes (the values are those of apply_edge, bit for bit): pass 1
// stores every face's nu * curl(v) and every node's divergence of v, pass 2 gathers them per edge
// (an edge reads four face terms and two node divergences instead of recomputing 16 + 12 edge
// values).
__device__ __forceinline__ uint32_t node_index(const EmGeo& g, const int p[3]) {
    return (uint32_t)p[0] + (uint32_t)(g.n[0] + 1) * ((uint32_t)p[1] + (uint32_t)(g.n[1] + 1) * (uint32_t)p[2]);
}
// The elements of one lattice (dims dx, dy, ...) split into B contiguous block ranges; a thread
// decodes its first element once and then steps by the block size, carrying x into y and z
// (no division per element).  f(x, y, z, k) with k the flat index in the lattice.
template <class F>
__device__ __forceinline__ void for_lattice(uint32_t dx, uint32_t dy, uint32_t total, uint32_t local, uint32_t B,
                                            F&& f) {
    uint32_t lo, hi;
    block_range(VGrid{local, B}, total, lo, hi);
    uint32_t k = lo + threadIdx.x;
    if (k >= hi) return;
    uint32_t x = k % dx;
    const uint32_t t = k / dx;
    uint32_t y = t % dy, z = t / dy;
    for (; k < hi; k += MHDP_BLOCK) {
        f((int)x, (int)y, (int)z, k);
        x += MHDP_BLOCK;
        while (x >= dx) {
            x -= dx;
            if (++y == dy) {
                y = 0;
                ++z;
            }
        }
    }
}
// pass 1: block group a < 3 the faces of axis a, group 3 the nodes (B blocks each)
struct K_em_curl {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        const float* nu;
        const float* v;
        float* tface;      // per face: nu * curl(v)
        float* ndiv;       // per node: node_div(v)
        const EmScal* sc;  // NULL: always run
        uint32_t B;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (a.sc && a.sc->status != 0) return;
        const EmGeo g = a.g;
        // the four groups interleaved: consecutive logical blocks (one XCD, one L2) take the same
        // slab of the three face lattices and of the nodes
        const uint32_t group = vg.b % 4u, local = vg.b / 4u;
        const uint32_t nx = (uint32_t)g.n[0], ny = (uint32_t)g.n[1], nz = (uint32_t)g.n[2];
        const float* __restrict__ v = a.v;
        if (group < 3) {
            const int axis = (int)group;
            const uint32_t off = axis == 0 ? 0u : (axis == 1 ? g.F[0] : g.F[0] + g.F[1]);
            float* __restrict__ t = a.tface + off;
            const uint32_t dx = nx + (axis == 0), dy = ny + (axis == 1), dz = nz + (axis == 2);
            const int ea = (axis + 1) % 3, eb = (axis + 2) % 3;
            for_lattice(dx, dy, dx * dy * dz, local, a.B, [&](int x, int y, int z, uint32_t k) {
                // face_term of a face in the lattice: nu * curl, the four edges loaded up front
                const int p[3] = {x, y, z};
                int qa[3] = {x, y, z}, qb[3] = {x, y, z};
                qa[ea] += 1;
                qb[eb] += 1;
                const bool v0 = edge_valid(g, eb, qa), v1 = edge_valid(g, eb, p), v2 = edge_valid(g, ea, qb),
                           v3 = edge_valid(g, ea, p);
                float e0 = v[v0 ? edge_index(g, eb, qa) : 0u], e1 = v[v1 ? edge_index(g, eb, p) : 0u];
                float e2 = v[v2 ? edge_index(g, ea, qb) : 0u], e3 = v[v3 ? edge_index(g, ea, p) : 0u];
                float n = a.nu[off + k];
                asm volatile("" : "+v"(e0), "+v"(e1), "+v"(e2), "+v"(e3), "+v"(n));
                const float cu = (v0 ? e0 : 0.0f) - (v1 ? e1 : 0.0f) - (v2 ? e2 : 0.0f) + (v3 ? e3 : 0.0f);
                t[k] = n * cu;
            });
        } else {
            float* __restrict__ d = a.ndiv;
            for_lattice(nx + 1, ny + 1, (nx + 1) * (ny + 1) * (nz + 1), local, a.B, [&](int x, int y, int z, uint32_t k) {
                // node_div with its six edges loaded up front
                const int p[3] = {x, y, z};
                bool ok[3][2];
                float e[3][2];
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    int q[3] = {x, y, z};
                    q[b] -= 1;
                    ok[b][0] = edge_valid(g, b, p);
                    ok[b][1] = edge_valid(g, b, q);
                    e[b][0] = v[ok[b][0] ? edge_index(g, b, p) : 0u];
                    e[b][1] = v[ok[b][1] ? edge_index(g, b, q) : 0u];
                }
#pragma unroll
                for (int b = 0; b < 3; b++) asm volatile("" : "+v"(e[b][0]), "+v"(e[b][1]));
                float value = 0.0f;
#pragma unroll
                for (int b = 0; b < 3; b++) value += (ok[b][0] ? e[b][0] : 0.0f) - (ok[b][1] ? e[b][1] : 0.0f);
                d[k] = value;
            });
        }
    }
};
__device__ __forceinline__ float stored_term(const EmGeo& g, const float* __restrict__ t, int axis, const int p[3]) {
    return face_valid(g, axis, p) ? t[face_index(g, axis, p)] : 0.0f;
}
// apply_operator from the stored face terms and node divergences
__device__ __forceinline__ float apply_stored_at(const EmGeo& g, const Phys& ph, const float* __restrict__ tface,
                                                 const float* __restrict__ ndiv, const uint8_t* __restrict__ flags,
                                                 const float* __restrict__ sigma, const float* __restrict__ v,
                                                 uint32_t i, int axis, const int p[3]) {
    // every value either form may need is loaded first, from indices clamped into their arrays
    // (a clamped value is never used: air edges are interior, invalid faces count as 0), and pinned
    // so that the loads stay ahead of the arithmetic - one memory round trip per edge
    uint32_t d[3];
    edge_dims(g, axis, d);
    const uint32_t zs = d[0] * d[1], last = g.Etot - 1u;
    const uint32_t im = i >= 1u ? i - 1u : i, ip = i + 1u <= last ? i + 1u : i;
    const uint32_t iym = i >= d[0] ? i - d[0] : i, iyp = i + d[0] <= last ? i + d[0] : i;
    const uint32_t izm = i >= zs ? i - zs : i, izp = i + zs <= last ? i + zs : i;
    const int ea = (axis + 1) % 3, eb = (axis + 2) % 3;
    int qa[3] = {p[0], p[1], p[2]}, qb[3] = {p[0], p[1], p[2]}, upper[3] = {p[0], p[1], p[2]};
    qa[ea] -= 1;
    qb[eb] -= 1;
    upper[axis] += 1;
    const bool f0 = face_valid(g, eb, p), f1 = face_valid(g, eb, qa), f2 = face_valid(g, ea, p), f3 = face_valid(g, ea, qb);
    uint32_t fl = flags[i];
    float c = v[i], xm = v[im], xp = v[ip], ym = v[iym], yp = v[iyp], zm = v[izm], zp = v[izp];
    float t0 = tface[f0 ? face_index(g, eb, p) : 0u], t1 = tface[f1 ? face_index(g, eb, qa) : 0u];
    float t2 = tface[f2 ? face_index(g, ea, p) : 0u], t3 = tface[f3 ? face_index(g, ea, qb) : 0u];
    float n0 = ndiv[node_index(g, p)], n1 = ndiv[node_index(g, upper)];
    float sg = sigma[i];
    asm volatile("" : "+v"(fl), "+v"(c), "+v"(xm), "+v"(xp), "+v"(ym), "+v"(yp), "+v"(zm), "+v"(zp));
    asm volatile("" : "+v"(t0), "+v"(t1), "+v"(t2), "+v"(t3), "+v"(n0), "+v"(n1), "+v"(sg));
    float q;
    if (fl & kScalarAir) {  // air_laplacian (induction.wgsl:150-158)
        float result = 0.0f;
        result += (c - xm) + (c - xp);
        result += (c - ym) + (c - yp);
        result += (c - zm) + (c - zp);
        q = result;
    } else {  // adjoint + gauge + mass (induction.wgsl:164)
        const float adj = (f0 ? t0 : 0.0f) - (f1 ? t1 : 0.0f) - (f2 ? t2 : 0.0f) + (f3 ? t3 : 0.0f);
        float gg = 0.0f;
        if (fl & kLowerAir) gg += n0;
        if (fl & kUpperAir) gg -= n1;
        q = adj + gg + mass(g, ph, sg) * c;
    }
    if (fl & kBoundary) q = c;
    return q;
}
__device__ __forceinline__ float apply_stored(const EmGeo& g, const Phys& ph, const float* __restrict__ tface,
                                              const float* __restrict__ ndiv, const uint8_t* __restrict__ flags,
                                              const float* __restrict__ sigma, const float* __restrict__ v,
                                              uint32_t i) {
    int p[3];
    const int axis = edge_decode(g, i, p);
    return apply_stored_at(g, ph, tface, ndiv, flags, sigma, v, i, axis, p);
}

struct K_em_apply {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        Vecs v;
        const float* tface;
        const float* ndiv;
        const EmScal* sc;
        double* part;
        uint32_t B;
    };
    // the edges of axis vg.b % 3, by rows
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (a.sc->status != 0) return;
        const EmGeo g = a.g;
        // the three axes interleaved: consecutive logical blocks (one XCD, one L2) take the same slab of
        // the three edge lattices, which share their face terms and node divergences
        const int axis = (int)(vg.b % 3u);
        const uint32_t local = vg.b / 3u;
        const uint32_t off = axis == 0 ? 0u : (axis == 1 ? g.E[0] : g.E[0] + g.E[1]);
        const float* __restrict__ pv = a.v.p;
        float* __restrict__ q = a.v.q;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const uint32_t dx = (uint32_t)g.n[0] + (axis != 0), dy = (uint32_t)g.n[1] + (axis != 1);
        for_lattice(dx, dy, g.E[axis], local, a.B, [&](int x, int y, int z, uint32_t k) {
                     const int p[3] = {x, y, z};
                     const uint32_t i = off + k;
                     const float qi = apply_stored_at(g, a.ph, a.tface, a.ndiv, a.v.flags, a.v.sigma, pv, i, axis, p);
                     q[i] = qi;
                     acc[0] += (double)pv[i] * (double)qi;
                 });
        block_partials<0, 0>(acc, a.part, vg.b);
    }
};

// update_residual (induction.wgsl:177-187): partials (r^2 / d, r^2)
struct K_em_update {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        Vecs v;
        const EmScal* sc;
        double* part;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (a.sc->status != 0) return;
        const float alpha = (float)a.sc->alpha, scale = (float)a.sc->scale;
        uint32_t lo, hi;
        block_range(vg, a.g.Etot, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const float r = a.v.r[i] - alpha * a.v.q[i];
            a.v.r[i] = r;
            a.v.x[i] = a.v.x[i] + scale * (alpha * a.v.p[i]);
            const float d = diagonal(a.g, a.ph, a.v.diag[i], a.v.sigma[i]);
            acc[0] += (double)r * (double)r / (double)d;
            acc[1] += (double)r * (double)r;
        }
        block_partials<0, 0>(acc, a.part, vg.b);
    }
};

// update_search (induction.wgsl:195): p = z + beta p, z = r / d (Jacobi) or the nodal
// correction's z, which the nodal pass left in q
struct K_em_search {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        Vecs v;
        const EmScal* sc;
        int nodal;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (a.sc->status != 0) return;
        const float beta = (float)a.sc->beta;
        uint32_t lo, hi;
        block_range(vg, a.g.Etot, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const float z = a.nodal ? a.v.q[i] : a.v.r[i] / diagonal(a.g, a.ph, a.v.diag[i], a.v.sigma[i]);
            a.v.p[i] = z + beta * a.v.p[i];
        }
    }
};

// ---- nodal auxiliary-space correction (nodal.wgsl): z = r / d + G S^-1 G^T r, G the node-to-edge
// incidence gradient, S = diag(G^T A G) on interior nodes in the reference's closed form (36 air(p)
// plus, over the three axes, the mass of the two incident edges and the air flags of the two
// neighbours).  The node values live in the operator's per-node scratch (ndiv: written by the
// curl pass, dead after the apply), so the correction allocates nothing.

// cached_air_node (nodal.wgsl:6): the lower-air flag of the node's x edge, 0 off the interior
__device__ __forceinline__ bool interior_node(const EmGeo& g, const int p[3]) {
    return p[0] > 0 && p[1] > 0 && p[2] > 0 && p[0] < g.n[0] && p[1] < g.n[1] && p[2] < g.n[2];
}
__device__ __forceinline__ float air_at(const EmGeo& g, const uint8_t* __restrict__ flags, const int p[3]) {
    if (!interior_node(g, p)) return 0.0f;
    return (flags[edge_index(g, 0, p)] & kLowerAir) ? 1.0f : 0.0f;
}

// nodal_build (nodal.wgsl:7-19): the node values N = S^-1 (r_out - r_in summed over the axes)
struct K_em_nodal_build {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        const float* r;
        const float* sigma;
        const uint8_t* flags;
        float* nodal;
        const EmScal* sc;  // NULL: always (nodal_build_initial)
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (a.sc && a.sc->status != 0) return;
        const EmGeo g = a.g;
        const uint32_t NX = (uint32_t)g.n[0] + 1u, NY = (uint32_t)g.n[1] + 1u;
        const uint32_t nodes = NX * NY * ((uint32_t)g.n[2] + 1u);
        uint32_t lo, hi;
        block_range(vg, nodes, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            const int p[3] = {(int)(i % NX), (int)((i / NX) % NY), (int)(i / (NX * NY))};
            float value = 0.0f;
            if (interior_node(g, p)) {
                float rhs = 0.0f, diag = 36.0f * air_at(g, a.flags, p);
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    int lower[3] = {p[0], p[1], p[2]}, upper[3] = {p[0], p[1], p[2]};
                    lower[ax] -= 1;
                    upper[ax] += 1;
                    const uint32_t out = edge_index(g, ax, p), inc = edge_index(g, ax, lower);
                    rhs += a.r[out] - a.r[inc];
                    diag += mass(g, a.ph, a.sigma[out]) + mass(g, a.ph, a.sigma[inc]) + air_at(g, a.flags, lower) +
                            air_at(g, a.flags, upper);
                }
                value = rhs / fmaxf(diag, 1e-30f);
            }
            a.nodal[i] = value;
        }
    }
};

// nodal_initial / nodal_update (nodal.wgsl:20-41): z = r / d + N(lower node) - N(upper node) into p
// (initial: the search starts at z) or q (the update's z for the next search); partials r.z, r.r
// replace those of the pass before (same blocks, same edges)
struct K_em_nodal_z {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        Vecs v;
        const float* nodal;
        const EmScal* sc;
        double* part;
        int initial;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        if (!a.initial && a.sc->status != 0) return;
        const EmGeo g = a.g;
        const uint32_t NX = (uint32_t)g.n[0] + 1u, NY = (uint32_t)g.n[1] + 1u;
        float* __restrict__ out = a.initial ? a.v.p : a.v.q;
        uint32_t lo, hi;
        block_range(vg, g.Etot, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int p[3];
            const int axis = edge_decode(g, i, p);
            int u[3] = {p[0], p[1], p[2]};
            u[axis] += 1;
            const uint32_t lower = (uint32_t)p[0] + NX * ((uint32_t)p[1] + NY * (uint32_t)p[2]);
            const uint32_t upper = (uint32_t)u[0] + NX * ((uint32_t)u[1] + NY * (uint32_t)u[2]);
            const float r = a.v.r[i];
            const float z = r / diagonal(g, a.ph, a.v.diag[i], a.v.sigma[i]) + a.nodal[lower] - a.nodal[upper];
            out[i] = z;
            acc[0] += (double)r * (double)z;
            acc[1] += (double)r * (double)r;
        }
        block_partials<0, 0, 2>(acc, a.part, vg.b);
    }
};

// true_residual (induction.wgsl:232-243): restart the search from the actual residual
struct K_em_true_residual {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        Vecs v;
        const EmScal* sc;
        double* part;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        const float scale = (float)a.sc->scale;
        uint32_t lo, hi;
        block_range(vg, a.g.Etot, lo, hi);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) {
            int p[3];
            const int axis = edge_decode(a.g, i, p);
            const float b = em_rhs(a.g, a.ph, a.v, i, axis, p, true) / scale;
            const float z = b / diagonal(a.g, a.ph, a.v.diag[i], a.v.sigma[i]);
            a.v.r[i] = b;
            a.v.p[i] = z;
            acc[0] += (double)b * (double)z;
            acc[1] += (double)b * (double)b;
        }
        block_partials<0, 0>(acc, a.part, vg.b);
    }
};

// commit_step: the new potential
struct K_em_commit {
    static constexpr int MINB = 0;
    struct Args {
        uint32_t n;
        const float* a_prev;
        const float* x;
        float* out;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        uint32_t lo, hi;
        block_range(vg, a.n, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK) a.out[i] = a.a_prev[i] + a.x[i];
    }
};

enum EmFold { kFoldInitial, kFoldAlpha, kFoldBeta, kFoldRestart, kFoldCheck };

// The single-block reductions (initial_reduce, alpha_reduce, beta_reduce, restart_reduce,
// coefficient_check; induction.wgsl:100-104, 130-148, 171-176, 189-194, 244-247)
struct K_em_fold {
    static constexpr int MINB = 0;
    struct Args {
        const double* part;
        uint32_t blocks;
        int mode;
        EmScal* sc;
    };
    static __device__ void run(const Args& a, const VGrid&) {
        __shared__ double red[4][MHDP_BLOCK];
        const bool check = a.mode == kFoldCheck;
        double v[4] = {0.0, check ? 1e30 : 0.0, 0.0, 0.0};
        for (uint32_t b = threadIdx.x; b < a.blocks; b += MHDP_BLOCK) {
            const double* w = a.part + 4u * b;
            v[0] += w[0];
            v[1] = check ? fmin(v[1], w[1]) : v[1] + w[1];
            v[2] = check ? fmax(v[2], w[2]) : v[2] + w[2];
            v[3] += w[3];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) red[k][threadIdx.x] = v[k];
        __syncthreads();
        for (int s = MHDP_BLOCK / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                const int t = threadIdx.x;
                red[0][t] += red[0][t + s];
                red[1][t] = check ? fmin(red[1][t], red[1][t + s]) : red[1][t] + red[1][t + s];
                red[2][t] = check ? fmax(red[2][t], red[2][t + s]) : red[2][t] + red[2][t + s];
                red[3][t] += red[3][t + s];
            }
            __syncthreads();
        }
        if (threadIdx.x != 0) return;
        EmScal& s = *a.sc;
        const double t0 = red[0][0], t1 = red[1][0], t2 = red[2][0], t3 = red[3][0];
        switch (a.mode) {
        case kFoldInitial: {
            // initial_reduce: an exact power-of-two scale when |b|_1 is tiny or huge
            const double scale = s.scale;
            double next = scale;
            const float l1 = (float)t3;
            if (scale == 1.0 && l1 > 0.0f && (l1 < 1e-8f || l1 > 1e8f)) {
                int ex = (int)((__float_as_uint(l1) >> 23) & 255u) - 127;
                ex = ex < -100 ? -100 : (ex > 100 ? 100 : ex);
                next = ldexp(1.0, ex);
            }
            double floor_sq = 1e-38;
            if (scale != 1.0) {
                const double fn = 1e-19 / scale;
                floor_sq = fn * fn;
            }
            s.rz = t0;
            s.rr = t1;
            s.bb = t2;
            s.l1 = t3;
            s.threshold = t2 * s.tol2 + floor_sq;
            s.status = t1 <= s.threshold ? 1 : 0;
            s.rescale = next != scale ? 1 : 0;
            s.scale = next;
            break;
        }
        case kFoldAlpha:
            if (s.status != 0) break;
            if (!(t0 > 0.0) || fabs(t0) > 3e38) s.status = 2;
            else s.alpha = s.rz / t0;
            break;
        case kFoldBeta:
            if (s.status != 0) break;
            s.beta = t0 / s.rz;
            s.rz = t0;
            s.rr = t1;
            s.its += 1;
            if (t1 <= s.threshold) s.status = 1;
            if (fabs(t0) > 3e38 || fabs(t1) > 3e38) s.status = 2;
            break;
        case kFoldRestart:
            s.rz = t0;
            s.rr = t1;
            s.status = t1 <= s.threshold ? 1 : 0;
            break;
        default:
            s.check[0] = t0;
            s.check[1] = t1;
            s.check[2] = t2;
            s.check[3] = t3;
            break;
        }
    }
};

// Blocks of the edge kernels: up to 2048 (measured at 128^3: the passes move 85-205 MB in 23-51 us,
// and twice the pressure kernels' cap gives the waves that hide their gathers' latency; 61.6 ->
// 53.2 ms per solve step)
uint32_t em_grid(uint32_t n) { return grid_blocks(n, 2048); }
// blocks per lattice of the row kernels: about a quarter of the edge kernels' grid each, a multiple
// of 8 (every XCD gets whole groups), so that 4 groups stay within the partial records
uint32_t row_blocks(const EmGeo& g) {
    const uint32_t b = em_grid(g.Etot) / 4 / 8 * 8;
    return b < 8 ? 8 : b;
}

EmGeo geo_of(const EmDims& d) {
    EmGeo g;
    for (int a = 0; a < 3; a++) g.n[a] = (int)d.n[a];
    for (int a = 0; a < 3; a++) {
        uint32_t e = 1, f = 1;
        for (int b = 0; b < 3; b++) {
            e *= d.n[b] + (b == a ? 0u : 1u);
            f *= d.n[b] + (b == a ? 1u : 0u);
        }
        g.E[a] = e;
        g.F[a] = f;
    }
    g.Etot = g.E[0] + g.E[1] + g.E[2];
    g.h = d.h;
    return g;
}

}  // namespace

uint32_t em_blocks(const EmDims& d) {
    const EmGeo g = geo_of(d);
    const uint32_t a = em_grid(g.Etot), b = 4 * row_blocks(g);
    return a > b ? a : b;
}

void launch_em_coefficients(hipStream_t s, const EmDims& d, const uint32_t* materials, const float* props, int vacuum,
                            float* sigma, float* diag, uint8_t* flags, float* nu, double* part, void* scal) {
    const EmGeo g = geo_of(d);
    const uint32_t faces = g.F[0] + g.F[1] + g.F[2];
    go<K_em_nu>(s, em_grid(faces), {g, materials, props, vacuum, nu});
    const uint32_t blocks = em_grid(g.Etot);
    go<K_em_coef>(s, blocks, {g, materials, props, vacuum, sigma, diag, flags, part});
    go<K_em_fold>(s, 1, {part, blocks, kFoldCheck, (EmScal*)scal});
}

void launch_em_begin(hipStream_t s, const EmDims& d, const uint8_t* flags, const float* a_prev, const float* a_before,
                     float prediction, float* x) {
    const EmGeo g = geo_of(d);
    go<K_em_begin>(s, em_grid(g.Etot), {g.Etot, flags, a_prev, a_before, prediction, x});
}

namespace {
// nodal_build_initial + nodal_initial after a pass that wrote r (prepare, true_residual) and its
// partials over em_grid(Etot) blocks: the search starts at the corrected z
void nodal_start(hipStream_t s, const EmGeo& g, const Phys& ph, const Vecs& v, double* part, EmScal* sc) {
    const uint32_t nodes = ((uint32_t)g.n[0] + 1u) * ((uint32_t)g.n[1] + 1u) * ((uint32_t)g.n[2] + 1u);
    go<K_em_nodal_build>(s, em_grid(nodes), {g, ph, v.r, v.sigma, v.flags, v.ndiv, nullptr});
    go<K_em_nodal_z>(s, em_grid(g.Etot), {g, ph, v, v.ndiv, sc, part, 1});
}
}  // namespace

void launch_em_prepare(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part, void* scal,
                       int guess, bool nodal) {
    const EmGeo g = geo_of(d);
    const uint32_t blocks = em_grid(g.Etot);
    const Phys& p = ph;
    const Vecs& vv = v;
    go<K_em_prepare>(s, blocks, {g, p, vv, (EmScal*)scal, part, guess});
    if (nodal) nodal_start(s, g, p, vv, part, (EmScal*)scal);
    go<K_em_fold>(s, 1, {part, blocks, kFoldInitial, (EmScal*)scal});
}

void launch_em_iterations(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part, void* scal,
                          uint32_t count, bool nodal, KernelTimer* timer) {
    KernelTimer off;
    KernelTimer& T = timer ? *timer : off;
    const EmGeo g = geo_of(d);
    const uint32_t blocks = em_grid(g.Etot);
    const Phys& p = ph;
    const Vecs& vv = v;
    EmScal* sc = (EmScal*)scal;
    const uint32_t B = row_blocks(g);
    for (uint32_t k = 0; k < count; k++) {
        T.before(0, s);
        go<K_em_curl>(s, 4 * B, {g, vv.nu, vv.p, v.tface, v.ndiv, sc, B});
        T.after(0, s);
        T.before(1, s);
        go<K_em_apply>(s, 3 * B, {g, p, vv, v.tface, v.ndiv, sc, part, B});
        T.after(1, s);
        go<K_em_fold>(s, 1, {part, 3 * B, kFoldAlpha, sc});
        T.before(2, s);
        go<K_em_update>(s, blocks, {g, p, vv, sc, part});
        T.after(2, s);
        if (nodal) {  // nodal_build + nodal_update: z of the new residual in q, its r.z in the partials
            const uint32_t nodes = ((uint32_t)g.n[0] + 1u) * ((uint32_t)g.n[1] + 1u) * ((uint32_t)g.n[2] + 1u);
            go<K_em_nodal_build>(s, em_grid(nodes), {g, p, vv.r, vv.sigma, vv.flags, vv.ndiv, sc});
            go<K_em_nodal_z>(s, blocks, {g, p, vv, vv.ndiv, sc, part, 0});
        }
        go<K_em_fold>(s, 1, {part, blocks, kFoldBeta, sc});
        T.before(3, s);
        go<K_em_search>(s, blocks, {g, p, vv, sc, nodal ? 1 : 0});
        T.after(3, s);
    }
}

void launch_em_restart(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part, void* scal,
                       bool nodal) {
    const EmGeo g = geo_of(d);
    const uint32_t blocks = em_grid(g.Etot);
    const Phys& p = ph;
    const Vecs& vv = v;
    go<K_em_true_residual>(s, blocks, {g, p, vv, (EmScal*)scal, part});
    if (nodal) nodal_start(s, g, p, vv, part, (EmScal*)scal);
    go<K_em_fold>(s, 1, {part, blocks, kFoldRestart, (EmScal*)scal});
}

// The nodal correction applied to v.r (test hook): z into v.p, node values into v.ndiv
void launch_em_probe_nodal(hipStream_t s, const EmDims& d, const EmPhys& ph, const EmVecs& v, double* part,
                           void* scal) {
    const EmGeo g = geo_of(d);
    nodal_start(s, g, ph, v, part, (EmScal*)scal);
}

void launch_em_commit(hipStream_t s, const EmDims& d, const float* a_prev, const float* x, float* out) {
    const EmGeo g = geo_of(d);
    go<K_em_commit>(s, em_grid(g.Etot), {g.Etot, a_prev, x, out});
}

// apply_operator alone (test hook): q = A p
struct K_em_probe {
    static constexpr int MINB = 0;
    struct Args {
        EmGeo g;
        Phys ph;
        const float* tface;
        const float* ndiv;
        const uint8_t* flags;
        const float* sigma;
        const float* p;
        float* q;
    };
    static __device__ void run(const Args& a, const VGrid& vg) {
        uint32_t lo, hi;
        block_range(vg, a.g.Etot, lo, hi);
        for (uint32_t i = lo + threadIdx.x; i < hi; i += MHDP_BLOCK)
            a.q[i] = apply_stored(a.g, a.ph, a.tface, a.ndiv, a.flags, a.sigma, a.p, i);
    }
};

void launch_em_probe_apply(hipStream_t s, const EmDims& d, const EmPhys& ph, const float* nu, const uint8_t* flags,
                           const float* sigma, const float* p, float* q, float* tface, float* ndiv) {
    const EmGeo g = geo_of(d);
    const uint32_t B = row_blocks(g);
    go<K_em_curl>(s, 4 * B, {g, nu, p, tface, ndiv, nullptr, B});
    go<K_em_probe>(s, em_grid(g.Etot), {g, ph, tface, ndiv, flags, sigma, p, q});
}
