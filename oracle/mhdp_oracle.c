/*
 * CPU oracle for the MAC pressure projection — TEST INFRASTRUCTURE ONLY
 * (checker and CPU baseline; never part of the product library).
 * Shared grid helpers; the precision-generic solver body lives in core.inc.
 */
#include "mhdp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    uint32_t d[3];
    uint64_t n;
    uint32_t periodic;
    const uint32_t *mask;
    const float *alpha;
    float h, rho_l, rho_air;
    int two_phase;
    int meanfree;
} ctx_t;

static void ctx_init(ctx_t *c, const orc_grid *g, const uint32_t *mask, const float *alpha, int meanfree) {
    c->d[0] = g->nx; c->d[1] = g->ny; c->d[2] = g->nz;
    c->n = (uint64_t)g->nx * g->ny * g->nz;
    c->periodic = g->periodic;
    c->mask = mask;
    c->alpha = alpha;
    c->h = g->h;
    c->rho_l = g->rho_liquid;
    c->rho_air = g->rho_air;
    c->two_phase = g->two_phase;
    c->meanfree = meanfree;
}

/* src/periodic.wgsl:5-10 — unsigned Euclidean wrap */
static inline int32_t wrap_periodic(int32_t value, uint32_t size) {
    if (value >= 0) return (int32_t)((uint32_t)value % size);
    uint32_t magnitude = (uint32_t)(-(value + 1)) + 1u;
    uint32_t remainder = magnitude % size;
    return (int32_t)(remainder == 0u ? 0u : size - remainder);
}
/* solver.wgsl:23-27 wrap_cell; multigrid.wgsl:18-20 canonical */
static inline void wrap_pos(uint32_t periodic, const uint32_t d[3], int32_t q[3]) {
    for (int a = 0; a < 3; a++)
        if (periodic & (1u << a)) q[a] = wrap_periodic(q[a], d[a]);
}
static inline int valid_pos(const uint32_t d[3], const int32_t q[3]) {
    return q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < (int32_t)d[0] && q[1] < (int32_t)d[1] &&
           q[2] < (int32_t)d[2];
}
static inline uint64_t idx3(const uint32_t d[3], const int32_t q[3]) {
    return (uint64_t)q[0] + (uint64_t)d[0] * ((uint64_t)q[1] + (uint64_t)d[1] * (uint64_t)q[2]);
}
static inline void decode3(uint64_t i, const uint32_t d[3], int32_t q[3]) {
    q[0] = (int32_t)(i % d[0]);
    q[1] = (int32_t)((i / d[0]) % d[1]);
    q[2] = (int32_t)(i / ((uint64_t)d[0] * d[1]));
}
static inline void plane_dims(const uint32_t d[3], int axis, uint32_t P[3]) {
    P[0] = d[0]; P[1] = d[1]; P[2] = d[2];
    P[axis] += 1;
}
static inline uint64_t plane_size(const uint32_t d[3], int axis) {
    uint32_t P[3];
    plane_dims(d, axis, P);
    return (uint64_t)P[0] * P[1] * P[2];
}
static inline uint64_t plane_offset(const uint32_t d[3], int axis) {
    uint64_t o = 0;
    for (int a = 0; a < axis; a++) o += plane_size(d, a);
    return o;
}
/* index of a canonical face position in its SoA plane */
static inline uint64_t face_idx(const uint32_t d[3], int axis, const int32_t q[3]) {
    uint32_t P[3];
    plane_dims(d, axis, P);
    return plane_offset(d, axis) + idx3(P, q);
}

uint64_t orc_face_plane_size(const orc_grid *g, int axis) {
    uint32_t d[3] = {g->nx, g->ny, g->nz};
    return plane_size(d, axis);
}

/* solver.wgsl:28-29 */
static inline int fluid(const ctx_t *c, const int32_t pos[3]) {
    int32_t q[3] = {pos[0], pos[1], pos[2]};
    wrap_pos(c->periodic, c->d, q);
    if (!valid_pos(c->d, q)) return 0;
    return c->mask[idx3(c->d, q)] != 0u;
}
/* solver.wgsl:31-32 */
static inline int valid_face(const ctx_t *c, const int32_t q[3], int axis) {
    uint32_t P[3];
    plane_dims(c->d, axis, P);
    return valid_pos(P, q);
}
static inline int open_face(const ctx_t *c, const int32_t pos[3], int axis) {
    int32_t q[3] = {pos[0], pos[1], pos[2]};
    wrap_pos(c->periodic, c->d, q);
    if (!valid_face(c, q, axis)) return 0;
    int32_t left[3] = {q[0], q[1], q[2]};
    left[axis] -= 1;
    return fluid(c, left) && fluid(c, q);
}
/* solver.wgsl:33-36 face(): value at the canonical face of an open face, else 0 */
static inline float face_value(const ctx_t *c, const float *planes, const int32_t pos[3], int axis) {
    int32_t q[3] = {pos[0], pos[1], pos[2]};
    wrap_pos(c->periodic, c->d, q);
    if (!open_face(c, q, axis)) return 0.0f;
    return planes[face_idx(c->d, axis, q)];
}
/* solver.wgsl:40-51 (mushy factor off: flags.w == 0) */
static inline float occupancy(const ctx_t *c, const int32_t pos[3]) {
    int32_t q[3] = {pos[0], pos[1], pos[2]};
    wrap_pos(c->periodic, c->d, q);
    if (!valid_pos(c->d, q)) return 0.0f;
    return c->alpha ? c->alpha[idx3(c->d, q)] : 0.0f;
}
static inline float density(const ctx_t *c, const int32_t pos[3]) {
    float a = occupancy(c, pos);
    return c->rho_air * (1.0f - a) + c->rho_l * a; /* mix(rho_air, rho_l, a) */
}
static inline float face_density(const ctx_t *c, const int32_t pos[3], int axis) {
    int32_t left[3] = {pos[0], pos[1], pos[2]};
    left[axis] -= 1;
    return 0.5f * (density(c, left) + density(c, pos));
}
static inline float mobility(const ctx_t *c, const int32_t pos[3], int axis) {
    float beta = 1.0f;
    if (c->two_phase) beta = c->rho_l / face_density(c, pos, axis);
    return beta;
}
/* solver.wgsl:60-64 */
static float diagonal(const ctx_t *c, const int32_t pos[3]) {
    float d = 0.0f;
    for (int axis = 0; axis < 3; axis++)
        for (int sign = -1; sign <= 1; sign += 2) {
            int32_t q[3] = {pos[0], pos[1], pos[2]};
            q[axis] += sign;
            if (fluid(c, q)) {
                int32_t fp[3] = {pos[0], pos[1], pos[2]};
                if (sign > 0) fp[axis] += 1;
                d += mobility(c, fp, axis);
            }
        }
    if (!c->meanfree && (c->mask[idx3(c->d, pos)] & 3u) == 2u) return 1.0f;
    return fmaxf(d, 1e-30f);
}

/* src/fluid/mod.rs:1447-1493 */
uint32_t orc_anchors(uint32_t *mask, const uint32_t dims[3], uint32_t periodic) {
    uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
    uint32_t components = 0;
    uint32_t *roots = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint64_t *queue = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    for (uint64_t root = 0; root < n; root++) {
        if (mask[root] != 1u) continue;
        components++;
        mask[root] = 4u;
        roots[root] = (uint32_t)root;
        uint64_t head = 0, tail = 0;
        queue[tail++] = root;
        while (head < tail) {
            uint64_t i = queue[head++];
            int32_t p[3];
            decode3(i, dims, p);
            for (int axis = 0; axis < 3; axis++)
                for (int dir = -1; dir <= 1; dir += 2) {
                    int32_t q[3] = {p[0], p[1], p[2]};
                    int32_t v = p[axis] + dir;
                    if (periodic & (1u << axis)) {
                        int32_t m = (int32_t)dims[axis];
                        v = ((v % m) + m) % m; /* rem_euclid */
                    } else if (v < 0 || v >= (int32_t)dims[axis]) {
                        continue;
                    }
                    q[axis] = v;
                    uint64_t j = idx3(dims, q);
                    if (mask[j] == 1u) {
                        mask[j] = 3u;
                        roots[j] = (uint32_t)root;
                        queue[tail++] = j;
                    }
                }
        }
    }
    for (uint64_t i = 0; i < n; i++) {
        if (mask[i] == 3u) mask[i] = 1u | (roots[i] << 2);
        else if (mask[i] == 4u) mask[i] = 2u | (roots[i] << 2);
    }
    free(roots);
    free(queue);
    return components;
}

/* src/fluid/multigrid.rs:25-65 */
int orc_hierarchy(const uint32_t dims[4], uint32_t *level_dims, int max_levels, uint64_t *coarse_cells) {
    uint64_t count = 1;
    for (int a = 0; a < 3; a++) {
        if (dims[a] > 0x7fffffffu) return -1;
        count *= dims[a];
        if (count > 0xffffffffull && a < 2) { /* checked_mul on u64 never overflows here */ }
    }
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0 || count != dims[3]) return -1;
    int levels = 0;
    uint32_t cur[4] = {dims[0], dims[1], dims[2], dims[3]};
    if (levels < max_levels) memcpy(level_dims + 4 * levels, cur, sizeof cur);
    levels++;
    uint64_t offset = 0;
    while (cur[0] > 4 || cur[1] > 4 || cur[2] > 4) {
        uint32_t nd[4];
        for (int a = 0; a < 3; a++) nd[a] = (cur[a] + 1) / 2;
        nd[3] = nd[0] * nd[1] * nd[2];
        if (offset > 0xffffffffull) return -1;
        if (levels < max_levels) memcpy(level_dims + 4 * levels, nd, sizeof nd);
        levels++;
        offset += nd[3];
        if (offset > 0xffffffffull) return -1;
        memcpy(cur, nd, sizeof nd);
    }
    if (coarse_cells) *coarse_cells = offset;
    return levels;
}

/* pressure_coefficients (solver.wgsl:93-103) over SoA face planes + per-cell diagonal */
void orc_coefficients(const orc_grid *g, const uint32_t *mask, const float *alpha, int meanfree,
                      float *beta, float *diag) {
    ctx_t c;
    ctx_init(&c, g, mask, alpha, meanfree);
    for (int axis = 0; axis < 3; axis++) {
        uint32_t P[3];
        plane_dims(c.d, axis, P);
        uint64_t np = (uint64_t)P[0] * P[1] * P[2];
        float *out = beta + plane_offset(c.d, axis);
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < (int64_t)np; f++) {
            int32_t pos[3];
            decode3((uint64_t)f, P, pos);
            out[f] = open_face(&c, pos, axis) ? mobility(&c, pos, axis) : 0.0f;
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)c.n; i++) {
        float d = 0.0f;
        if (mask[i] != 0u) {
            int32_t pos[3];
            decode3((uint64_t)i, c.d, pos);
            d = diagonal(&c, pos);
        }
        diag[i] = d;
    }
}

void orc_diagonal_uncached(const orc_grid *g, const uint32_t *mask, const float *alpha, int meanfree,
                           float *diag) {
    ctx_t c;
    ctx_init(&c, g, mask, alpha, meanfree);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)c.n; i++) {
        int32_t pos[3];
        decode3((uint64_t)i, c.d, pos);
        diag[i] = mask[i] ? diagonal(&c, pos) : 0.0f;
    }
}

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ viscous stress (two-phase)
 * diffuse_surface_stress (src/fluid/mod.rs:1146-1234), src/fluid/viscosity.wgsl and the RKC2
 * stages of src/fluid/stabilized.{rs,wgsl}.  The packet's reserved.z is the occupancy. */
static inline float stress_mu(const ctx_t *c, float nu_rho, float mu_air, const int32_t q[3]) {
    if (!fluid(c, q)) return 0.0f; /* dynamic_mu (viscosity.wgsl:3-6) */
    const float t = fminf(fmaxf(occupancy(c, q), 0.0f), 1.0f);
    return mu_air * (1.0f - t) + nu_rho * t;
}
static float stress_edge_mu(const ctx_t *c, float nu_rho, float mu_air, const int32_t q[3], int i, int j) {
    float total = 0.0f, count = 0.0f; /* coefficient_edge_mu (viscosity.wgsl:16-22) */
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            int32_t at[3] = {q[0], q[1], q[2]};
            at[i] -= a;
            at[j] -= b;
            if (fluid(c, at)) {
                total += stress_mu(c, nu_rho, mu_air, at);
                count += 1.0f;
            }
        }
    return total / fmaxf(count, 1.0f);
}
static float normal_stress(const ctx_t *c, float nu_rho, float mu_air, const float *u, const int32_t q[3], int i) {
    if (!fluid(c, q)) return 0.0f; /* viscosity.wgsl:27-30 */
    int32_t upper[3] = {q[0], q[1], q[2]};
    upper[i] += 1;
    return 2.0f * stress_mu(c, nu_rho, mu_air, q) * (face_value(c, u, upper, i) - face_value(c, u, q, i)) / c->h;
}
static float shear_stress(const ctx_t *c, float nu_rho, float mu_air, const float *u, const int32_t q[3], int i, int j) {
    int32_t li[3] = {q[0], q[1], q[2]}, lj[3] = {q[0], q[1], q[2]}; /* viscosity.wgsl:31-42 */
    li[j] -= 1;
    lj[i] -= 1;
    float ui1 = face_value(c, u, q, i), ui0 = face_value(c, u, li, i);
    float uj1 = face_value(c, u, q, j), uj0 = face_value(c, u, lj, j);
    /* odd tangential ghosts: zero velocity halfway to a closed wall */
    if (!open_face(c, q, i) && open_face(c, li, i)) ui1 = -ui0;
    if (!open_face(c, li, i) && open_face(c, q, i)) ui0 = -ui1;
    if (!open_face(c, q, j) && open_face(c, lj, j)) uj1 = -uj0;
    if (!open_face(c, lj, j) && open_face(c, q, j)) uj0 = -uj1;
    return stress_edge_mu(c, nu_rho, mu_air, q, i, j) * (ui1 - ui0 + uj1 - uj0) / c->h;
}
/* stress_acceleration (viscosity.wgsl:68-82): div(tau) / (h rho_face) */
static float stress_acc(const ctx_t *c, float nu_rho, float mu_air, const float *u, const int32_t q[3], int i) {
    int32_t left[3] = {q[0], q[1], q[2]};
    left[i] -= 1;
    float div = normal_stress(c, nu_rho, mu_air, u, q, i) - normal_stress(c, nu_rho, mu_air, u, left, i);
    for (int j = 0; j < 3; j++) {
        if (j == i) continue;
        int32_t upper[3] = {q[0], q[1], q[2]};
        upper[j] += 1;
        div += shear_stress(c, nu_rho, mu_air, u, upper, i, j) - shear_stress(c, nu_rho, mu_air, u, q, i, j);
    }
    return div / (c->h * face_density(c, q, i));
}

/* the cached form the RKC2 stages use (viscosity.wgsl:56-76, tensor_normal / tensor_shear): shear
 * stresses come from the per-cell tensor, which holds each edge's stress once with the component
 * pair in ascending order (xy, xz, yz); positions outside the grid compute them directly. */
static float shear_cached(const ctx_t *c, float nu_rho, float mu_air, const float *u, const int32_t q[3], int i, int j) {
    int32_t at[3] = {q[0], q[1], q[2]};
    wrap_pos(c->periodic, c->d, at);
    if (!valid_pos(c->d, at)) return shear_stress(c, nu_rho, mu_air, u, q, i, j);
    return i < j ? shear_stress(c, nu_rho, mu_air, u, at, i, j) : shear_stress(c, nu_rho, mu_air, u, at, j, i);
}
static float stress_acc_cached(const ctx_t *c, float nu_rho, float mu_air, const float *u, const int32_t q[3], int i) {
    int32_t left[3] = {q[0], q[1], q[2]};
    left[i] -= 1;
    float div = normal_stress(c, nu_rho, mu_air, u, q, i) - normal_stress(c, nu_rho, mu_air, u, left, i);
    for (int j = 0; j < 3; j++) {
        if (j == i) continue;
        int32_t upper[3] = {q[0], q[1], q[2]};
        upper[j] += 1;
        div += shear_cached(c, nu_rho, mu_air, u, upper, i, j) - shear_cached(c, nu_rho, mu_air, u, q, i, j);
    }
    return div / (c->h * face_density(c, q, i));
}

static float *packet_alpha(const orc_grid *g, const float *cells) {
    const uint64_t n = (uint64_t)g->nx * g->ny * g->nz;
    float *alpha = (float *)malloc((n + 1) * sizeof(float));
    for (uint64_t i = 0; i < n; i++) alpha[i] = cells[16 * i + 14];
    return alpha;
}

/* viscosity_bound + reduce (viscosity.wgsl:83-104): the largest row sum of rho^-1 div(tau) */
float orc_stress_bound(const orc_grid *g, const uint32_t *mask, const float *cells, float nu, float mu_air) {
    float *alpha = packet_alpha(g, cells);
    ctx_t c;
    ctx_init(&c, g, mask, alpha, 0);
    const float nu_rho = nu * g->rho_liquid;
    float bound = 0.0f;
    for (int axis = 0; axis < 3; axis++) {
        uint32_t P[3];
        plane_dims(c.d, axis, P);
        const uint64_t np = plane_size(c.d, axis);
        for (uint64_t k = 0; k < np; k++) {
            int32_t q[3];
            decode3(k, P, q);
            if (!open_face(&c, q, axis)) continue;
            int32_t left[3] = {q[0], q[1], q[2]};
            left[axis] -= 1;
            float sum = stress_mu(&c, nu_rho, mu_air, q) + stress_mu(&c, nu_rho, mu_air, left);
            for (int j = 0; j < 3; j++) {
                if (j == axis) continue;
                int32_t upper[3] = {q[0], q[1], q[2]};
                upper[j] += 1;
                sum += stress_edge_mu(&c, nu_rho, mu_air, q, axis, j) + stress_edge_mu(&c, nu_rho, mu_air, upper, axis, j);
            }
            bound = fmaxf(bound, 4.0f * sum / (face_density(&c, q, axis) * c.h * c.h));
        }
    }
    free(alpha);
    return bound;
}

/* one explicit substep, stress_diffuse (viscosity.wgsl:105-112): out = u + dt_sub div(tau)/rho */
void orc_stress_euler(const orc_grid *g, const uint32_t *mask, const float *cells, float nu, float mu_air,
                      float dt_sub, const float *u, float *out) {
    float *alpha = packet_alpha(g, cells);
    ctx_t c;
    ctx_init(&c, g, mask, alpha, 0);
    const float nu_rho = nu * g->rho_liquid;
    for (int axis = 0; axis < 3; axis++) {
        uint32_t P[3];
        plane_dims(c.d, axis, P);
        const uint64_t off = plane_offset(c.d, axis), np = plane_size(c.d, axis);
#pragma omp parallel for schedule(static)
        for (uint64_t k = 0; k < np; k++) {
            int32_t q[3];
            decode3(k, P, q);
            float v = 0.0f;
            if (open_face(&c, q, axis))
                v = face_value(&c, u, q, axis) + dt_sub * stress_acc(&c, nu_rho, mu_air, u, q, axis);
            out[off + k] = v;
        }
    }
    free(alpha);
}

/* RKC2 stage rows (stabilized.rs:13-53): damped second-order Runge-Kutta-Chebyshev (Verwer,
 * Hundsdorfer & Sommeijer 1990) with damping 0.05 and the fewest stages whose stability interval
 * covers 1.8 x euler_steps.  rows[8 k]: (mu, nu, 1 - mu - nu, dt-weighted operator, initial-operator
 * weight, 0, 0, 0).  Returns the number of stages. */
int orc_rkc_schedule(double dt, uint32_t euler_steps, float *rows, int max_rows) {
    const double damping = 0.05, stiffness = 1.8 * (double)euler_steps;
    const int capacity = (int)ceil(3.0 * sqrt((double)euler_steps)) + 2;
    for (int s = 2; s <= capacity; s++) {
        const double w0 = 1.0 + damping / ((double)s * s);
        double t[512], d[512], dd[512], b[512], a[512];
        if (s + 1 > 512 || s > max_rows) return -1;
        t[0] = 1.0; t[1] = w0; d[0] = 0.0; d[1] = 1.0; dd[0] = 0.0; dd[1] = 0.0;
        for (int j = 2; j <= s; j++) {
            t[j] = 2.0 * w0 * t[j - 1] - t[j - 2];
            d[j] = 2.0 * t[j - 1] + 2.0 * w0 * d[j - 1] - d[j - 2];
            dd[j] = 4.0 * d[j - 1] + 2.0 * w0 * dd[j - 1] - dd[j - 2];
        }
        const double w1 = d[s] / dd[s];
        if (stiffness > 0.9 * (1.0 + w0) / w1) continue;
        for (int j = 2; j <= s; j++) b[j] = dd[j] / (d[j] * d[j]);
        b[0] = b[2];
        b[1] = b[2];
        for (int j = 0; j <= s; j++) a[j] = 1.0 - b[j] * t[j];
        float *r = rows;
        memset(rows, 0, sizeof(float) * 8 * (size_t)s);
        r[0] = 1.0f;
        r[3] = (float)(dt * b[1] * w1);
        for (int j = 2; j <= s; j++) {
            float *row = rows + 8 * (j - 1);
            const float mu = (float)(2.0 * w0 * b[j] / b[j - 1]);
            const float nu = (float)(-b[j] / b[j - 2]);
            const double op = dt * 2.0 * w1 * b[j] / b[j - 1];
            row[0] = mu;
            row[1] = nu;
            row[2] = 1.0f - mu - nu;
            row[3] = (float)op;
            row[4] = (float)(-a[j - 1] * op);
        }
        return s;
    }
    return -1;
}

/* stress_stabilized stage (stabilized.wgsl:18-30) */
static void rkc_stage(const ctx_t *c, float nu_rho, float mu_air, const float *row, const float *prev,
                      const float *older, const float *y0, const float *f0, float *out /* may be older */) {
    for (int axis = 0; axis < 3; axis++) {
        uint32_t P[3];
        plane_dims(c->d, axis, P);
        const uint64_t off = plane_offset(c->d, axis), np = plane_size(c->d, axis);
#pragma omp parallel for schedule(static)
        for (uint64_t k = 0; k < np; k++) {
            int32_t q[3];
            decode3(k, P, q);
            float v = 0.0f;
            if (open_face(c, q, axis)) {
                int32_t w[3] = {q[0], q[1], q[2]};
                wrap_pos(c->periodic, c->d, w);
                const uint64_t canon = face_idx(c->d, axis, w);
                v = row[0] * face_value(c, prev, q, axis) + row[1] * older[off + k] + row[2] * y0[canon] +
                    row[3] * stress_acc_cached(c, nu_rho, mu_air, prev, q, axis) + row[4] * f0[canon];
            }
            out[off + k] = v;
        }
    }
}

/* diffuse_surface_stress (mod.rs:1160-1234) on u (three planes, in/out): returns 0, 4 (too many
 * substeps for max_substeps) or 3 (non-finite bound); *substeps as the reference's diagnostics. */
int orc_surface_stress(const orc_grid *g, const uint32_t *mask, const float *cells, double dt, float nu, float mu_air,
                       uint32_t max_substeps, float *u, uint32_t *substeps, double *count_out) {
    *substeps = 0;
    const double liquid = g->rho_liquid, ambient = g->rho_air;
    const double eps = 1.1920928955078125e-07;
    const double slack = (double)1e-5f + 4.0 * eps;
    const double floor_rho = fmin(liquid, ambient) - fabs(liquid - ambient) * slack;
    double count;
    int proved = 0;
    if (floor_rho > 0.0) {
        const double mu = (double)fmaxf(nu * g->rho_liquid, mu_air);
        const double h = g->h;
        const double gb = 24.0 * mu / (floor_rho * h * h) * (1.0 + 32.0 * eps);
        if (dt * gb <= 1.8) {
            if (gb == 0.0) return 0;
            count = 1.0;
            proved = 1;
        }
    }
    if (!proved) {
        const float bound = orc_stress_bound(g, mask, cells, nu, mu_air);
        if (!isfinite(bound) || bound < 0.0f) return 3;
        if (bound == 0.0f) return 0;
        count = ceil(dt * (double)bound / 1.8);
        if (count < 1.0) count = 1.0;
    }
    if (count_out) *count_out = count;
    if (count > (double)max_substeps) return 4;
    const uint64_t F = plane_offset((uint32_t[3]){g->nx, g->ny, g->nz}, 3);
    float *tmp = (float *)malloc((F + 1) * sizeof(float));
    const uint32_t n = (uint32_t)count;
    if (n > 8) {
        float rows[8 * 256];
        const int stages = orc_rkc_schedule(dt, n, rows, 256);
        float *alpha = packet_alpha(g, cells);
        ctx_t c;
        ctx_init(&c, g, mask, alpha, 0);
        const float nu_rho = nu * g->rho_liquid;
        float *y0 = (float *)malloc((F + 1) * sizeof(float)), *f0 = (float *)malloc((F + 1) * sizeof(float));
        float *older = (float *)calloc(F + 1, sizeof(float));
        memcpy(y0, u, F * sizeof(float));
        for (int axis = 0; axis < 3; axis++) { /* stress_save_initial */
            uint32_t P[3];
            plane_dims(c.d, axis, P);
            const uint64_t off = plane_offset(c.d, axis), np = plane_size(c.d, axis);
            for (uint64_t k = 0; k < np; k++) {
                int32_t q[3];
                decode3(k, P, q);
                f0[off + k] = open_face(&c, q, axis) ? stress_acc_cached(&c, nu_rho, mu_air, u, q, axis) : 0.0f;
            }
        }
        /* stage k reads Y_k (newest, at neighbours) and Y_{k-1} (its own element only) and writes
         * Y_{k+1} over Y_{k-1}: two buffers alternate (stabilized.rs:234-247); Y_{-1} = 0 */
        float *newest = u, *other = older;
        for (int st = 0; st < stages; st++) {
            rkc_stage(&c, nu_rho, mu_air, rows + 8 * st, newest, other, y0, f0, other);
            float *t2 = newest;
            newest = other;
            other = t2;
        }
        if (newest != u) memcpy(u, newest, F * sizeof(float));
        *substeps = (uint32_t)stages;
        free(y0); free(f0); free(alpha); free(older);
    } else {
        const float dt_sub = (float)(dt / (double)n);
        for (uint32_t k = 0; k < n; k++) {
            orc_stress_euler(g, mask, cells, nu, mu_air, dt_sub, u, tmp);
            memcpy(u, tmp, F * sizeof(float));
        }
        *substeps = n;
    }
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------------ EM edge lattice
 * src/em/common.wgsl, src/em/grid.rs: edges of axis a have dims n + 1 except n_a on axis a,
 * packed x | y | z; faces of axis a have dims n + e_a. */
enum { EM_BOUNDARY = 1, EM_LOWER_AIR = 2, EM_UPPER_AIR = 4, EM_SCALAR_AIR = 8 };
typedef struct {
    uint32_t n[3], E[3], Etot, F[3];
    float h, dt, mu0;
    int magnetostatic;
    const float *sigma, *diag, *nu;
    const uint8_t *flags;
} em_ctx;
static void em_ctx_init(em_ctx *c, const orc_em_grid *g, const orc_em_params *prm, const float *sigma,
                        const float *diag, const uint8_t *flags, const float *nu) {
    c->n[0] = g->nx; c->n[1] = g->ny; c->n[2] = g->nz;
    for (int a = 0; a < 3; a++) {
        uint32_t e = 1, f = 1;
        for (int b = 0; b < 3; b++) {
            e *= c->n[b] + (b == a ? 0u : 1u);
            f *= c->n[b] + (b == a ? 1u : 0u);
        }
        c->E[a] = e;
        c->F[a] = f;
    }
    c->Etot = c->E[0] + c->E[1] + c->E[2];
    c->h = g->h;
    c->dt = prm ? prm->dt : 1.0f;
    c->mu0 = prm ? prm->mu0 : 0.0f;
    c->magnetostatic = prm ? prm->magnetostatic : 1;
    c->sigma = sigma; c->diag = diag; c->flags = flags; c->nu = nu;
}
static inline void em_edge_dims(const em_ctx *c, int axis, uint32_t d[3]) {
    for (int b = 0; b < 3; b++) d[b] = c->n[b] + (b == axis ? 0u : 1u);
}
static inline int em_edge_valid(const em_ctx *c, int axis, const int32_t p[3]) {
    uint32_t d[3];
    em_edge_dims(c, axis, d);
    return p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && (uint32_t)p[0] < d[0] && (uint32_t)p[1] < d[1] && (uint32_t)p[2] < d[2];
}
static inline uint32_t em_edge_index(const em_ctx *c, int axis, const int32_t p[3]) {
    uint32_t d[3];
    em_edge_dims(c, axis, d);
    uint32_t off = 0;
    for (int a = 0; a < axis; a++) off += c->E[a];
    return off + (uint32_t)p[0] + d[0] * ((uint32_t)p[1] + d[1] * (uint32_t)p[2]);
}
static inline void em_edge_decode(const em_ctx *c, uint32_t i, int *axis, int32_t p[3]) {
    int a = 0;
    while (a < 2 && i >= c->E[a]) i -= c->E[a++];
    uint32_t d[3];
    em_edge_dims(c, a, d);
    p[0] = (int32_t)(i % d[0]);
    p[1] = (int32_t)((i / d[0]) % d[1]);
    p[2] = (int32_t)(i / (d[0] * d[1]));
    *axis = a;
}
static inline int em_face_valid(const em_ctx *c, int axis, const int32_t p[3]) {
    return p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && (uint32_t)p[0] < c->n[0] + (axis == 0) &&
           (uint32_t)p[1] < c->n[1] + (axis == 1) && (uint32_t)p[2] < c->n[2] + (axis == 2);
}
static inline uint32_t em_face_index(const em_ctx *c, int axis, const int32_t p[3]) {
    uint32_t off = 0;
    for (int a = 0; a < axis; a++) off += c->F[a];
    const uint32_t px = c->n[0] + (axis == 0), py = c->n[1] + (axis == 1);
    return off + (uint32_t)p[0] + px * ((uint32_t)p[1] + py * (uint32_t)p[2]);
}
/* boundary (common.wgsl:9): an edge on the outer surface (tangential A is held) */
static inline int em_boundary(const em_ctx *c, int axis, const int32_t p[3]) {
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    return p[a] == 0 || (uint32_t)p[a] == c->n[a] || p[b] == 0 || (uint32_t)p[b] == c->n[b];
}
/* cell_properties (induction.wgsl:16-25): (conductivity, relative permeability) of a cell;
 * props[2 k] / props[2 k + 1] for material id k; partially filled cells (bit 31) blend with id 0 */
static inline void em_cell(const em_ctx *c, const uint32_t *materials, const float *props, int vacuum,
                           const int32_t p[3], float out[2]) {
    if (vacuum || p[0] < 0 || p[1] < 0 || p[2] < 0 || (uint32_t)p[0] >= c->n[0] || (uint32_t)p[1] >= c->n[1] ||
        (uint32_t)p[2] >= c->n[2]) {
        out[0] = 0.0f;
        out[1] = 1.0f;
        return;
    }
    const uint32_t id = materials[(uint32_t)p[0] + c->n[0] * ((uint32_t)p[1] + c->n[1] * (uint32_t)p[2])];
    if (id == 0u) {
        out[0] = props[0];
        out[1] = props[1];
        return;
    }
    if (id & 0x80000000u) {
        const float alpha = (float)((id >> 8) & 0x7fffffu) / 8388607.0f;
        const uint32_t k = id & 255u;
        out[0] = alpha * props[2 * k];
        out[1] = props[1] * (1.0f - alpha) + props[2 * k + 1] * alpha; /* mix(mu_0, mu_k, alpha) */
        return;
    }
    out[0] = props[2 * id];
    out[1] = props[2 * id + 1];
}
/* nu_face (induction.wgsl:26-30) */
static float em_nu_face(const em_ctx *c, const uint32_t *materials, const float *props, int vacuum, int axis,
                        const int32_t p[3]) {
    if (!em_face_valid(c, axis, p)) return 0.0f;
    int32_t q[3] = {p[0], p[1], p[2]};
    q[axis] -= 1;
    float a[2], b[2];
    em_cell(c, materials, props, vacuum, p, a);
    em_cell(c, materials, props, vacuum, q, b);
    return 0.5f / a[1] + 0.5f / b[1];
}
/* air_node (induction.wgsl:58-63) */
static int em_air_node(const em_ctx *c, const uint32_t *materials, const float *props, int vacuum, const int32_t p[3]) {
    for (int a = 0; a < 3; a++)
        if (p[a] <= 0 || (uint32_t)p[a] >= c->n[a]) return 0;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int z = 0; z < 2; z++) {
                const int32_t q[3] = {p[0] - x, p[1] - y, p[2] - z};
                float v[2];
                em_cell(c, materials, props, vacuum, q, v);
                if (v[0] > 0.0f) return 0;
            }
    return 1;
}
/* coefficients + coefficient_check (induction.wgsl:82-104) */
void orc_em_coefficients(const orc_em_grid *g, const uint32_t *materials, const float *props, int vacuum,
                         float *sigma, float *diag, uint8_t *flags, float *nu, float check[4]) {
    em_ctx c;
    em_ctx_init(&c, g, NULL, NULL, NULL, NULL, NULL);
    for (int a = 0; a < 3; a++) {
        uint32_t off = 0;
        for (int b = 0; b < a; b++) off += c.F[b];
        const uint32_t px = c.n[0] + (a == 0), py = c.n[1] + (a == 1);
        for (uint32_t k = 0; k < c.F[a]; k++) {
            const int32_t p[3] = {(int32_t)(k % px), (int32_t)((k / px) % py), (int32_t)(k / (px * py))};
            nu[off + k] = em_nu_face(&c, materials, props, vacuum, a, p);
        }
    }
    check[0] = 0.0f; check[1] = 1e30f; check[2] = 0.0f; check[3] = 0.0f;
    for (uint32_t i = 0; i < c.Etot; i++) {
        int axis;
        int32_t p[3];
        em_edge_decode(&c, i, &axis, p);
        const int a = (axis + 1) % 3, b = (axis + 2) % 3;
        float s = 0.0f;
        for (int u = 0; u < 2; u++)
            for (int v = 0; v < 2; v++) {
                int32_t q[3] = {p[0], p[1], p[2]};
                q[a] -= u;
                q[b] -= v;
                float cp[2];
                em_cell(&c, materials, props, vacuum, q, cp);
                s += 0.25f * cp[0];
            }
        int32_t qa[3] = {p[0], p[1], p[2]}, qb[3] = {p[0], p[1], p[2]};
        qa[a] -= 1;
        qb[b] -= 1;
        const float n1 = em_nu_face(&c, materials, props, vacuum, b, p), n2 = em_nu_face(&c, materials, props, vacuum, b, qa);
        const float n3 = em_nu_face(&c, materials, props, vacuum, a, p), n4 = em_nu_face(&c, materials, props, vacuum, a, qb);
        float d = n1 + n2 + n3 + n4;
        uint8_t fl = 0;
        if (em_boundary(&c, axis, p)) {
            d = 1.0f;
            fl |= EM_BOUNDARY;
            if (s > 0.0f) check[0] += 1.0f;
        }
        int32_t upper[3] = {p[0], p[1], p[2]};
        upper[axis] += 1;
        const int la = em_air_node(&c, materials, props, vacuum, p), ua = em_air_node(&c, materials, props, vacuum, upper);
        if (la) fl |= EM_LOWER_AIR;
        if (ua) fl |= EM_UPPER_AIR;
        if (la && ua && n1 == 1.0f && n2 == 1.0f && n3 == 1.0f && n4 == 1.0f) fl |= EM_SCALAR_AIR;
        sigma[i] = s;
        diag[i] = d + (la ? 1.0f : 0.0f) + (ua ? 1.0f : 0.0f);
        flags[i] = fl;
        if (s > 0.0f) {
            check[1] = fminf(check[1], s);
            check[3] += 1.0f;
        }
        check[2] = fmaxf(check[2], s);
    }
}

#define R float
#define F(name) name##_f32
#define RC(x) x##f
#include "core.inc"
#undef R
#undef F
#undef RC

#define R double
#define F(name) name##_f64
#define RC(x) x
#include "core.inc"
#undef R
#undef F
#undef RC

#define R float
#define F(name) name##_f32
#include "em.inc"
#undef R
#undef F
#define R double
#define F(name) name##_f64
#include "em.inc"
#undef R
#undef F

double orc_true_residual(const orc_grid *g, const uint32_t *mask, const float *alpha,
                         const float *ustar, const double *phi) {
    ctx_t c;
    ctx_init(&c, g, mask, alpha, 0);
    double num = 0.0, den = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : num, den)
    for (int64_t i = 0; i < (int64_t)c.n; i++) {
        if (!mask[i]) continue;
        int32_t pos[3];
        decode3((uint64_t)i, c.d, pos);
        double b = rhs_cell_f64(&c, ustar, pos, 0);
        double r = b - laplacian_cell_f64(&c, phi, pos);
        num += r * r;
        den += b * b;
    }
    return den > 0.0 ? sqrt(num / den) : 0.0;
}

void orc_correct_f64(const orc_grid *g, const uint32_t *mask, const float *alpha, const float *ustar,
                     const double *phi, float *u_out) {
    ctx_t c;
    ctx_init(&c, g, mask, alpha, 0);
    correct_f64(&c, ustar, phi, u_out);
}

/* fields() (solver.wgsl:256-276) per cell in its FP32 operation order; totals as diagnostics()
 * (solver.wgsl:284-291) reduces them, here in FP64 and cell order. */
void orc_fields(const orc_grid *g, const uint32_t *mask, const float *faces, const double *phi, const float *em,
                const orc_fields_params *p, float *cells, double *totals) {
    ctx_t c;
    ctx_init(&c, g, mask, NULL, p->meanfree);
    for (int k = 0; k < 10; k++) totals[k] = 0.0;
    const float h = g->h, rho_l = g->rho_liquid;
    for (uint64_t i = 0; i < c.n; i++) {
        float *cell = cells + 16 * i;
        const float old[4] = {cell[12], cell[13], cell[14], cell[15]};
        float out[16];
        memset(out, 0, sizeof out);
        if (g->two_phase) { out[13] = old[1]; out[14] = old[2]; out[15] = old[3]; }
        if (mask[i]) {
            int32_t pos[3];
            decode3(i, c.d, pos);
            float v[3], fe = 0.0f, mf = 0.0f, sum = 0.0f;
            for (int a = 0; a < 3; a++) {
                int32_t q[3] = {pos[0], pos[1], pos[2]};
                q[a] += 1;
                const float lo = face_value(&c, faces, pos, a), hi = face_value(&c, faces, q, a);
                v[a] = 0.5f * (lo + hi);
                fe += 0.5f * (lo * lo + hi * hi);
                mf = fmaxf(mf, fmaxf(fabsf(lo), fabsf(hi)));
            }
            for (int a = 0; a < 3; a++) {  /* divergence(pos, false), solver.wgsl:104-106 */
                int32_t q[3] = {pos[0], pos[1], pos[2]};
                q[a] += 1;
                sum += face_value(&c, faces, q, a) - face_value(&c, faces, pos, a);
            }
            const float div = sum / h;
            float melt = p->mushy ? old[1] : 1.0f;
            if (g->two_phase) melt *= old[2];
            const float speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const float pv = p->meanfree ? (float)phi[i] - (float)phi[mask[i] >> 2] : (float)phi[i];
            const float pressure = pv * rho_l / p->dt;
            float force[3] = {0.0f, 0.0f, 0.0f};
            if (em) for (int a = 0; a < 3; a++) force[a] = em[16 * i + 8 + a];
            for (int a = 0; a < 3; a++) out[a] = v[a];
            out[3] = speed;
            out[4] = pressure;
            out[5] = div;
            for (int a = 0; a < 3; a++) out[8 + a] = rho_l * p->acceleration[a] + force[a];
            out[12] = 1.0f;
            out[13] = p->mushy ? old[1] : 1.0f;
            out[14] = old[2];
            out[15] = old[3];
            int bad = 0;
            for (int a = 0; a < 3; a++) bad |= v[a] != v[a] || force[a] != force[a] || fabsf(force[a]) > 1e30f;
            bad |= speed > 1e25f || !(fabsf(pressure) < 1e30f) || !(fabsf(div) < 1e30f);
            const float rho = g->two_phase ? g->rho_air * (1.0f - old[2]) + rho_l * old[2] : rho_l;
            const float ke = 0.5f * rho * fe * h * h * h;
            const float fl = sqrtf(force[0] * force[0] + force[1] * force[1] + force[2] * force[2]);
            totals[0] = fmax(totals[0], speed);
            totals[1] = fmax(totals[1], fabsf(div));
            totals[2] += (double)(div * div);
            totals[3] += (double)ke;
            totals[4] = fmax(totals[4], fabsf(pressure));
            totals[5] += (double)div;
            totals[6] = fmax(totals[6], fl);
            totals[7] += bad ? 1.0 : 0.0;
            totals[8] = fmax(totals[8], mf);
            totals[9] += (double)melt;
        }
        memcpy(cell, out, sizeof out);
    }
}

/* Float -> integer cell coordinate for the samplers: the (integral) value clamped to +-2^30 first,
 * NaN giving -2^30 (fmaxf returns the number), so that base + 1 cannot overflow; such positions lie
 * far outside any grid.  (The reference converts with WGSL's f32 -> i32, whose result for values
 * out of range is not pinned by any of its tests.) */
static inline int32_t cell_coord(float v) { return (int32_t)fminf(fmaxf(v, -1073741824.0f), 1073741824.0f); }
/* sample_component (solver.wgsl:109-116): trilinear interpolation of face component `axis` */
static float sample_component(const ctx_t *c, const float *faces, const float position[3], int axis) {
    int32_t base[3];
    float frac[3];
    for (int b = 0; b < 3; b++) {
        const float q = position[b] - (b == axis ? 0.0f : 0.5f);
        const float fl = floorf(q);
        base[b] = cell_coord(fl);
        frac[b] = q - fl; /* fract() */
    }
    float value = 0.0f;
    for (int z = 0; z < 2; z++)
        for (int y = 0; y < 2; y++)
            for (int x = 0; x < 2; x++) {
                const int o[3] = {x, y, z};
                float w[3];
                for (int b = 0; b < 3; b++) w[b] = o[b] > 0 ? frac[b] : 1.0f - frac[b];
                int32_t q[3] = {base[0] + x, base[1] + y, base[2] + z};
                value += face_value(c, faces, q, axis) * w[0] * w[1] * w[2];
            }
    return value;
}
static void sample_velocity(const ctx_t *c, const float *faces, const float position[3], float v[3]) {
    for (int a = 0; a < 3; a++) v[a] = sample_component(c, faces, position, a);
}

void orc_predict(const orc_grid *g, const uint32_t *mask, const float *faces, const orc_step_params *p, float *out) {
    ctx_t c;
    ctx_init(&c, g, mask, NULL, 0);
    const float s_half = 0.5f * p->dt / g->h, s_full = p->dt / g->h;
    for (int axis = 0; axis < 3; axis++) {
        uint32_t P[3];
        plane_dims(c.d, axis, P);
        const uint64_t off = plane_offset(c.d, axis), np = plane_size(c.d, axis);
#pragma omp parallel for schedule(static)
        for (uint64_t k = 0; k < np; k++) {
            int32_t pos[3];
            decode3(k, P, pos);
            float result = 0.0f;
            if (open_face(&c, pos, axis)) {
                result = face_value(&c, faces, pos, axis);
                if (p->advect) {
                    float at[3], v[3], mid[3], dep[3];
                    for (int b = 0; b < 3; b++) at[b] = (float)pos[b] + 0.5f;
                    at[axis] -= 0.5f;
                    sample_velocity(&c, faces, at, v);
                    for (int b = 0; b < 3; b++) mid[b] = at[b] - v[b] * s_half;
                    sample_velocity(&c, faces, mid, v);
                    for (int b = 0; b < 3; b++) dep[b] = at[b] - v[b] * s_full;
                    result = sample_component(&c, faces, dep, axis);
                }
            }
            out[off + k] = result;
        }
    }
}

/* em_force (solver.wgsl:108) */
static inline float em_component(const ctx_t *c, const float *em, const int32_t pos[3], int axis) {
    int32_t q[3] = {pos[0], pos[1], pos[2]};
    wrap_pos(c->periodic, c->d, q);
    if (!em || !valid_pos(c->d, q)) return 0.0f;
    return em[16 * idx3(c->d, q) + 8 + axis];
}
/* reserved.w of a cell (the interface curvature) */
static inline float curvature(const ctx_t *c, const float *cells, const int32_t pos[3]) {
    int32_t q[3] = {pos[0], pos[1], pos[2]};
    wrap_pos(c->periodic, c->d, q);
    return cells[16 * idx3(c->d, q) + 15];
}
/* capillary_force (solver.wgsl:52-57) */
static float capillary_force(const ctx_t *c, const float *cells, float sigma, const int32_t pos[3], int axis) {
    if (!c->two_phase || !open_face(c, pos, axis)) return 0.0f;
    int32_t left[3] = {pos[0], pos[1], pos[2]};
    left[axis] -= 1;
    const float al = occupancy(c, left), ar = occupancy(c, pos);
    const float kl = curvature(c, cells, left), kr = curvature(c, cells, pos);
    const float wl = al * (1.0f - al), wr = ar * (1.0f - ar);
    const float k = wl + wr > 1e-6f ? (wl * kl + wr * kr) / fmaxf(wl + wr, 1e-20f) : 0.5f * (kl + kr);
    return sigma * k * (ar - al) / c->h;
}

void orc_viscous(const orc_grid *g, const uint32_t *mask, const float *pred, const float *cells, const float *em,
                 const orc_step_params *p, float *out) {
    /* occupancy() reads reserved.z of the packet (solver.wgsl:40) */
    float *alpha = NULL;
    if (g->two_phase) {
        alpha = (float *)malloc(((size_t)g->nx * g->ny * g->nz + 1) * sizeof(float));
        for (uint64_t i = 0; i < (uint64_t)g->nx * g->ny * g->nz; i++) alpha[i] = cells[16 * i + 14];
    }
    ctx_t c;
    ctx_init(&c, g, mask, alpha, 0);
    const float h = g->h, nu_dt_h2 = p->viscosity * p->dt / (h * h);
    for (int axis = 0; axis < 3; axis++) {
        uint32_t P[3];
        plane_dims(c.d, axis, P);
        const uint64_t off = plane_offset(c.d, axis), np = plane_size(c.d, axis);
#pragma omp parallel for schedule(static)
        for (uint64_t k = 0; k < np; k++) {
            int32_t pos[3];
            decode3(k, P, pos);
            float result = 0.0f;
            if (open_face(&c, pos, axis)) {
                const float center = face_value(&c, pred, pos, axis);
                float laplace = 0.0f;
                if (!c.two_phase)
                    for (int d = 0; d < 3; d++)
                        for (int sign = -1; sign <= 1; sign += 2) {
                            int32_t q[3] = {pos[0], pos[1], pos[2]};
                            q[d] += sign;
                            float nb = face_value(&c, pred, q, axis);
                            if (!open_face(&c, q, axis) && d != axis) nb = -center; /* no-slip wall */
                            laplace += nb - center;
                        }
                int32_t left[3] = {pos[0], pos[1], pos[2]};
                left[axis] -= 1;
                float rho = c.rho_l, capillary = 0.0f;
                if (c.two_phase) {
                    rho = face_density(&c, pos, axis);
                    capillary = capillary_force(&c, cells, p->surface_tension, pos, axis) / rho;
                }
                const float acc = p->acceleration[axis] +
                                  (em_component(&c, em, left, axis) + em_component(&c, em, pos, axis)) / (2.0f * rho) +
                                  capillary;
                float beta = mobility(&c, pos, axis);
                if (c.two_phase) beta *= rho / c.rho_l;
                result = (center + nu_dt_h2 * laplace + p->dt * acc) * beta;
            }
            out[off + k] = result;
        }
    }
    free(alpha);
}

int orc_project(const orc_grid *g, const uint32_t *mask, const float *alpha, const float *ustar,
                const orc_opts *o, double *phi_out, float *u_out, orc_stats *st) {
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (o->threads > 0) omp_set_num_threads(o->threads);
#endif
    int rc = o->precision == 64 ? project_f64(g, mask, alpha, ustar, o, phi_out, u_out, st)
                                : project_f32(g, mask, alpha, ustar, o, phi_out, u_out, st);
#ifdef _OPENMP
    if (o->threads > 0) omp_set_num_threads(saved);
#endif
    return rc;
}

struct orc_mg {
    int precision;
    void *impl;
};

orc_mg *orc_mg_new(const orc_grid *g, const uint32_t *mask, const float *alpha, int meanfree, int precision) {
    orc_mg *m = (orc_mg *)calloc(1, sizeof(orc_mg));
    m->precision = precision;
    m->impl = precision == 64 ? (void *)probe_new_f64(g, mask, alpha, meanfree)
                              : (void *)probe_new_f32(g, mask, alpha, meanfree);
    return m;
}
void orc_mg_free(orc_mg *m) {
    if (!m) return;
    if (m->precision == 64) probe_free_f64(m->impl);
    else probe_free_f32(m->impl);
    free(m);
}
int orc_mg_levels(const orc_mg *m, uint32_t *dims4) {
    return m->precision == 64 ? probe_levels_f64(m->impl, dims4) : probe_levels_f32(m->impl, dims4);
}
void orc_mg_action(orc_mg *m, int level, const double *v, double *out) {
    if (m->precision == 64) probe_action_f64(m->impl, level, v, out);
    else probe_action_f32(m->impl, level, v, out);
}
void orc_mg_vcycle(orc_mg *m, const double *r, double *z) {
    if (m->precision == 64) probe_vcycle_f64(m->impl, r, z);
    else probe_vcycle_f32(m->impl, r, z);
}
void orc_mg_coarse_solve(orc_mg *m, const double *rhs, double *z) {
    if (m->precision == 64) probe_coarse_f64(m->impl, rhs, z, -1);
    else probe_coarse_f32(m->impl, rhs, z, -1);
}
void orc_mg_coarse_sweeps(orc_mg *m, const double *rhs, double *z, int sweeps) {
    if (m->precision == 64) probe_coarse_f64(m->impl, rhs, z, sweeps);
    else probe_coarse_f32(m->impl, rhs, z, sweeps);
}
void orc_mg_graph(const orc_mg *m, int level, double *coeff4, double *diag) {
    if (m->precision == 64) probe_graph_f64(m->impl, level, coeff4, diag);
    else probe_graph_f32(m->impl, level, coeff4, diag);
}
