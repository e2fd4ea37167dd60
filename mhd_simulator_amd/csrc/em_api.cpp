// C ABI of the EM step's Jacobi-CG (include/mhdp.h, mhdp_em_*): EmState's coefficient set-up
// (src/em/mod.rs:557-571) and the solve loop of EmState::advance_to (src/em/mod.rs:863-1010)
// over the kernels of em.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "mhdp.h"
#include "mhdp_kernels.h"

struct mhdp_em {
    EmDims d{};
    uint64_t cells = 0, edges = 0, faces = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool ready = false;  // materials set
    float sigma_range[2] = {0.0f, 0.0f};
    uint32_t* materials = nullptr;
    float* props = nullptr;
    float *sigma = nullptr, *diag = nullptr, *nu = nullptr;
    uint8_t* flags = nullptr;
    float *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
    float *tface = nullptr, *ndiv = nullptr;  // operator scratch: per face, per node
    double* part = nullptr;
    EmScal* sc = nullptr;
    EmScal* sc_h = nullptr;  // pinned host copy
    KernelTimer timer;       // mhdp_em_set_kernel_timing
    double timer_ms = 0.0;
    uint32_t timer_launches = 0;
};

namespace {

int fail(mhdp_em* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}

#define EM_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) return fail(e, MHDP_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

int alloc(mhdp_em* e, void** p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(e, MHDP_NO_MEMORY, "EM allocation exceeds the available device memory");
    }
    return MHDP_OK;
}

int read_scalars(mhdp_em* e) {
    EM_HIP(hipMemcpyAsync(e->sc_h, e->sc, sizeof(EmScal), hipMemcpyDeviceToHost, e->stream));
    EM_HIP(hipStreamSynchronize(e->stream));
    EM_HIP(hipGetLastError());
    return MHDP_OK;
}

bool finite_status(const EmScal& s) {
    return std::isfinite(s.rz) && std::isfinite(s.rr) && std::isfinite(s.bb) && std::isfinite(s.threshold);
}

}  // namespace

extern "C" {

int mhdp_em_create(mhdp_em** out, const mhdp_grid* grid) {
    if (!out || !grid) return MHDP_INVALID;
    *out = nullptr;
    if (grid->nx == 0 || grid->ny == 0 || grid->nz == 0 || grid->periodic != 0 || !(grid->h > 0.0f) ||
        !std::isfinite(grid->h))
        return MHDP_INVALID;
    const uint64_t x = grid->nx, y = grid->ny, z = grid->nz;
    const uint64_t edges = x * (y + 1) * (z + 1) + (x + 1) * y * (z + 1) + (x + 1) * (y + 1) * z;
    if (edges > UINT32_MAX || x * y * z > UINT32_MAX) return MHDP_INVALID;  // "EM grid indexing exceeds u32"
    mhdp_em* e = new mhdp_em;
    e->d.n[0] = grid->nx;
    e->d.n[1] = grid->ny;
    e->d.n[2] = grid->nz;
    e->d.h = grid->h;
    e->cells = x * y * z;
    e->edges = edges;
    e->faces = (x + 1) * y * z + x * (y + 1) * z + x * y * (z + 1);
    const size_t E = edges;
    int rc;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&e->sc_h, sizeof(EmScal)) != hipSuccess) {
        mhdp_em_destroy(e);
        return MHDP_DEVICE;
    }
    if ((rc = alloc(e, (void**)&e->materials, e->cells * 4)) || (rc = alloc(e, (void**)&e->props, 256 * 2 * 4)) ||
        (rc = alloc(e, (void**)&e->sigma, E * 4)) || (rc = alloc(e, (void**)&e->diag, E * 4)) ||
        (rc = alloc(e, (void**)&e->nu, e->faces * 4)) || (rc = alloc(e, (void**)&e->flags, E)) ||
        (rc = alloc(e, (void**)&e->x, E * 4)) || (rc = alloc(e, (void**)&e->r, E * 4)) ||
        (rc = alloc(e, (void**)&e->p, E * 4)) || (rc = alloc(e, (void**)&e->q, E * 4)) ||
        (rc = alloc(e, (void**)&e->tface, e->faces * 4)) || (rc = alloc(e, (void**)&e->ndiv, (x + 1) * (y + 1) * (z + 1) * 4)) ||
        (rc = alloc(e, (void**)&e->part, (size_t)em_blocks(e->d) * 4 * sizeof(double) + 64)) ||
        (rc = alloc(e, (void**)&e->sc, sizeof(EmScal)))) {
        mhdp_em_destroy(e);
        return rc;
    }
    *out = e;
    return MHDP_OK;
}

int mhdp_em_set_materials(mhdp_em* e, const uint32_t* materials, const float* props, uint32_t count, int vacuum,
                          float sigma_range[2]) {
    if (!e) return MHDP_INVALID;
    e->err.clear();
    if (!materials || !props || count == 0 || count > 256) return fail(e, MHDP_INVALID, "Invalid EM material table");
    for (uint64_t i = 0; i < e->cells; i++) {
        const uint32_t id = materials[i];
        const uint32_t k = (id & 0x80000000u) ? (id & 255u) : id;
        if (k >= count) return fail(e, MHDP_INVALID, "EM material id outside the material table");
    }
    for (uint32_t k = 0; k < 2 * count; k++)
        if (!std::isfinite(props[k]) || (k % 2 == 1 && !(props[k] > 0.0f)) || (k % 2 == 0 && props[k] < 0.0f))
            return fail(e, MHDP_INVALID, "Invalid EM material properties");
    e->ready = false;
    EM_HIP(hipMemcpyAsync(e->materials, materials, e->cells * 4, hipMemcpyHostToDevice, e->stream));
    EM_HIP(hipMemcpyAsync(e->props, props, (size_t)count * 2 * 4, hipMemcpyHostToDevice, e->stream));
    launch_em_coefficients(e->stream, e->d, e->materials, e->props, vacuum ? 1 : 0, e->sigma, e->diag, e->flags, e->nu,
                           e->part, e->sc);
    int rc = read_scalars(e);
    if (rc) return rc;
    const double* check = e->sc_h->check;
    e->sigma_range[0] = check[3] > 0.0 ? (float)check[1] : 0.0f;
    e->sigma_range[1] = (float)check[2];
    if (sigma_range) {
        sigma_range[0] = e->sigma_range[0];
        sigma_range[1] = e->sigma_range[1];
    }
    if (check[0] > 0.0)  // src/em/mod.rs:566-570
        return fail(e, MHDP_INVALID, "A conductor touches the EM outer boundary; add air padding to the domain");
    e->ready = true;
    return MHDP_OK;
}

int mhdp_em_solve_dev(mhdp_em* e, const mhdp_em_opts* o, const float* as_now, const float* as_prev,
                      const float* motional, const float* a_prev, float* a_out, mhdp_em_stats* st) {
    return mhdp_em_solve_predicted_dev(e, o, as_now, as_prev, motional, a_prev, nullptr, 0.0f, a_out, st);
}

int mhdp_em_solve_predicted_dev(mhdp_em* e, const mhdp_em_opts* o, const float* as_now, const float* as_prev,
                                const float* motional, const float* a_prev, const float* a_before, float prediction,
                                float* a_out, mhdp_em_stats* st) {
    if (!e) return MHDP_INVALID;
    e->err.clear();
    if (!e->ready) return fail(e, MHDP_INVALID, "EM materials are not set");
    if (!o || !as_now || !as_prev || !a_prev || !a_out) return fail(e, MHDP_INVALID, "Missing EM solve argument");
    if (!(o->dt > 0.0f) || !std::isfinite(o->dt) || !(o->mu0 > 0.0f) || !(o->rel_tol > 0.0f) ||
        !(o->rel_tol < 1.0f) || o->max_iter == 0 || o->nodal < -1 || o->nodal > 1)
        return fail(e, MHDP_INVALID, "Invalid EM solve options");
    // the nodal correction (mod.rs:887-897): induction solves only, when the cap exceeds the switch
    // point; from the start when that point is 0
    const uint32_t switch_after = o->nodal == 1 ? o->nodal_after : 1024u;
    const bool allow = !o->magnetostatic && o->nodal >= 0 && o->max_iter > switch_after;
    bool nodal = allow && switch_after == 0;
    const auto t0 = std::chrono::steady_clock::now();
    const EmPhys ph{o->dt, o->mu0, o->magnetostatic ? 1 : 0};
    const EmVecs v{e->x,     e->r,    e->p,    e->q,     e->sigma, e->diag,  e->nu,
                   e->flags, as_now,  as_prev, motional, a_prev,   e->tface, e->ndiv};
    // begin_step: scale 1; the increment starts from prediction * (a_prev - a_before) (the
    // reference predicts after a regular step with a positive finite ratio, mod.rs:869-886) or
    // from zero; prepare + initial_reduce, repeated once on a new scale
    const bool guess = a_before && !o->magnetostatic && std::isfinite(prediction) && prediction > 0.0f;
    EmScal init{};
    init.scale = 1.0;
    init.tol2 = (double)(o->rel_tol * o->rel_tol);  // Params.physical.w: f32 relative_tolerance.powi(2)
    EM_HIP(hipMemcpyAsync(e->sc, &init, sizeof init, hipMemcpyHostToDevice, e->stream));
    if (guess) launch_em_begin(e->stream, e->d, e->flags, a_prev, a_before, prediction, e->x);
    int rc;
    for (int pass = 0; pass < 2; pass++) {
        launch_em_prepare(e->stream, e->d, ph, v, e->part, e->sc, guess ? 1 : 0, nodal);
        if ((rc = read_scalars(e))) return rc;
        if (!e->sc_h->rescale) break;
    }
    const double start_residual = std::sqrt(e->sc_h->rr / std::max(e->sc_h->bb, 1e-38));
    bool dropped = false;
    if (guess && (!finite_status(*e->sc_h) || !std::isfinite(e->sc_h->l1) || e->sc_h->rr > e->sc_h->bb)) {
        // a bad extrapolation must not start farther from the same right-hand side: zero_guess
        // and prepare again (mod.rs:906-925); the scale picked above stays
        launch_em_prepare(e->stream, e->d, ph, v, e->part, e->sc, 0, nodal);
        if ((rc = read_scalars(e))) return rc;
        dropped = true;
    }
    const uint32_t max_iter = o->max_iter;
    e->timer.used = 0;
    EmScal s = *e->sc_h;
    for (;;) {
        while (s.status == 0 && finite_status(s) && s.its < max_iter) {
            const uint32_t before = s.its;
            launch_em_iterations(e->stream, e->d, ph, v, e->part, e->sc, std::min<uint32_t>(max_iter - s.its, 32u),
                                 nodal, &e->timer);
            if ((rc = read_scalars(e))) return rc;
            s = *e->sc_h;
            // every batch either counts iterations or ends the solve (status); anything else is a
            // device-side fault of the reductions, reported instead of waiting for ever
            if (s.status == 0 && s.its == before) return fail(e, MHDP_DEVICE, "EM solve made no progress");
            if (allow && !nodal && s.status == 0 && finite_status(s) && s.its >= switch_after && s.its < max_iter) {
                // switch at a true-residual restart: the right-hand side norm, the acceptance
                // threshold and the iteration count carry over (mod.rs:941-959)
                nodal = true;
                launch_em_restart(e->stream, e->d, ph, v, e->part, e->sc, true);
                if ((rc = read_scalars(e))) return rc;
                s = *e->sc_h;
            }
        }
        if (s.status != 1) break;
        // converged on the recursive residual: verify with the actual equation and restart from it
        launch_em_restart(e->stream, e->d, ph, v, e->part, e->sc, nodal);
        if ((rc = read_scalars(e))) return rc;
        s = *e->sc_h;
        if (s.status == 1 || s.its >= max_iter) break;
    }
    if (s.status != 1 && finite_status(s)) {
        launch_em_restart(e->stream, e->d, ph, v, e->part, e->sc, nodal);
        if ((rc = read_scalars(e))) return rc;
        s = *e->sc_h;
    }
    // timed launches that did work (longer than a tenth of the longest; the rest returned at once)
    e->timer_ms = 0.0;
    e->timer_launches = 0;
    {
        std::vector<float> dur;
        for (size_t k = 0; k + 1 < e->timer.used; k += 2) {
            float t = 0.0f;
            EM_HIP(hipEventElapsedTime(&t, e->timer.ev[k], e->timer.ev[k + 1]));
            dur.push_back(t);
        }
        float longest = 0.0f;
        for (float t : dur) longest = std::max(longest, t);
        for (float t : dur)
            if (t >= 0.1f * longest) {
                e->timer_ms += t;
                e->timer_launches++;
            }
    }
    const double rel = std::sqrt(s.rr / std::max(s.bb, 1e-38));
    if (st) {
        st->iterations = s.its;
        st->rel_residual = rel;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        st->prediction = guess ? prediction : 0.0f;
        st->start_residual = start_residual;
        st->prediction_dropped = dropped ? 1u : 0u;
        st->nodal = nodal ? 1u : 0u;
    }
    if (s.status != 1 || !finite_status(s)) {  // src/em/mod.rs:1011-1021
        char buf[320];
        std::snprintf(buf, sizeof buf,
                      "EM solve did not converge: %u iterations, true relative residual %.3e (requested %.3e); "
                      "conducting-edge \xcf\x83 range %.3e\xe2\x80\xa6%.3e S/m; state restored",
                      s.its, rel, (double)o->rel_tol, (double)e->sigma_range[0], (double)e->sigma_range[1]);
        return fail(e, finite_status(s) ? MHDP_NOT_CONVERGED : MHDP_NONFINITE, buf);
    }
    launch_em_commit(e->stream, e->d, a_prev, e->x, a_out);  // commit_step
    EM_HIP(hipGetLastError());
    EM_HIP(hipStreamSynchronize(e->stream));
    return MHDP_OK;
}

int mhdp_em_get_coefficients(const mhdp_em* e, float* sigma, float* diag, uint8_t* flags, float* nu) {
    if (!e || !e->ready) return MHDP_INVALID;
    const size_t E = e->edges;
    if ((sigma && hipMemcpyAsync(sigma, e->sigma, E * 4, hipMemcpyDeviceToHost, e->stream) != hipSuccess) ||
        (diag && hipMemcpyAsync(diag, e->diag, E * 4, hipMemcpyDeviceToHost, e->stream) != hipSuccess) ||
        (flags && hipMemcpyAsync(flags, e->flags, E, hipMemcpyDeviceToHost, e->stream) != hipSuccess) ||
        (nu && hipMemcpyAsync(nu, e->nu, e->faces * 4, hipMemcpyDeviceToHost, e->stream) != hipSuccess) ||
        hipStreamSynchronize(e->stream) != hipSuccess)
        return MHDP_DEVICE;
    return MHDP_OK;
}

int mhdp_em_probe_apply(mhdp_em* e, const mhdp_em_opts* o, const float* p, float* q) {
    if (!e || !o || !p || !q) return MHDP_INVALID;
    e->err.clear();
    if (!e->ready) return fail(e, MHDP_INVALID, "EM materials are not set");
    const size_t bytes = e->edges * 4;
    EM_HIP(hipMemcpyAsync(e->p, p, bytes, hipMemcpyHostToDevice, e->stream));
    launch_em_probe_apply(e->stream, e->d, EmPhys{o->dt, o->mu0, o->magnetostatic ? 1 : 0}, e->nu, e->flags, e->sigma,
                          e->p, e->q, e->tface, e->ndiv);
    EM_HIP(hipMemcpyAsync(q, e->q, bytes, hipMemcpyDeviceToHost, e->stream));
    EM_HIP(hipStreamSynchronize(e->stream));
    EM_HIP(hipGetLastError());
    return MHDP_OK;
}

int mhdp_em_probe_nodal(mhdp_em* e, const mhdp_em_opts* o, const float* r, float* z, float* nodal) {
    if (!e || !o || !r || !z) return MHDP_INVALID;
    e->err.clear();
    if (!e->ready) return fail(e, MHDP_INVALID, "EM materials are not set");
    const size_t bytes = e->edges * 4;
    const EmPhys ph{o->dt, o->mu0, o->magnetostatic ? 1 : 0};
    const EmVecs v{e->x,     e->r,    e->p,    e->q,    e->sigma, e->diag,  e->nu,
                   e->flags, nullptr, nullptr, nullptr, nullptr,  e->tface, e->ndiv};
    EM_HIP(hipMemcpyAsync(e->r, r, bytes, hipMemcpyHostToDevice, e->stream));
    launch_em_probe_nodal(e->stream, e->d, ph, v, e->part, e->sc);
    EM_HIP(hipMemcpyAsync(z, e->p, bytes, hipMemcpyDeviceToHost, e->stream));
    if (nodal)
        EM_HIP(hipMemcpyAsync(nodal, e->ndiv, (size_t)(e->d.n[0] + 1) * (e->d.n[1] + 1) * (e->d.n[2] + 1) * 4,
                              hipMemcpyDeviceToHost, e->stream));
    EM_HIP(hipStreamSynchronize(e->stream));
    EM_HIP(hipGetLastError());
    return MHDP_OK;
}

uint64_t mhdp_em_edges(const mhdp_em* e) { return e ? e->edges : 0; }

int mhdp_em_set_kernel_timing(mhdp_em* e, int kernel) {
    if (!e || kernel < MHDP_EM_K_NONE || kernel > MHDP_EM_K_SEARCH) return MHDP_INVALID;
    e->timer.kernel = kernel;
    return MHDP_OK;
}

int mhdp_em_kernel_time(const mhdp_em* e, double* total_ms, uint32_t* launches) {
    if (!e) return MHDP_INVALID;
    if (total_ms) *total_ms = e->timer_ms;
    if (launches) *launches = e->timer_launches;
    return MHDP_OK;
}

const char* mhdp_em_last_error(const mhdp_em* e) { return e ? e->err.c_str() : "invalid EM context"; }

void mhdp_em_destroy(mhdp_em* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    void* ptrs[] = {e->materials, e->props, e->sigma, e->diag, e->nu,    e->flags, e->x,
                    e->r,         e->p,     e->q,     e->tface, e->ndiv, e->part,  e->sc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (e->sc_h) (void)hipHostFree(e->sc_h);
    for (hipEvent_t ev : e->timer.ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

}  // extern "C"
