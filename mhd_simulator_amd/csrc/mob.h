// Face mobilities formed from the liquid fraction, equal bit for bit to the coefficient pass's
// stored ones: density() and mobility() (solver.wgsl:40-51) in the same FP32 operations, each
// rounded on its own (the including files may contract multiply-adds elsewhere; these may not).
// Which faces are open comes from the per-cell open bytes: bit a set when the cell's lower face on
// axis a is open (the cell and its lower, wrapped neighbour fluid), bit 3 fluid, bit 4 the anchor
// of its component.  A cell is active (carries an unknown) when fluid and, in the pinned gauge,
// not an anchor.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct MobP {
    float rho_l, rho_air;
};
__device__ __forceinline__ float mob_dens(const MobP& m, float a) {
#pragma clang fp contract(off)
    return m.rho_air * (1.0f - a) + m.rho_l * a;
}
// mobility of a face between cells of densities dl and dc: 0 closed, 1 single phase (TWO false)
template <bool TWO>
__device__ __forceinline__ float mob_face(const MobP& m, bool open, float dl, float dc) {
#pragma clang fp contract(off)
    if (!open) return 0.0f;
    if (!TWO) return 1.0f;
    const float fd = 0.5f * (dl + dc);
    return m.rho_l / fd;
}
__device__ __forceinline__ uint32_t obyte(uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; }
__device__ __forceinline__ bool open_active(uint32_t bits, bool meanfree) {
    return (bits & 8u) && (meanfree || !(bits & 16u));
}
