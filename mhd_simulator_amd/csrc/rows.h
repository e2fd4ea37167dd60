// Row-vector helpers shared by the cell-parallel kernels that own four consecutive cells of
// a row (nx % 4 == 0): neighbour-row starts with periodic wrap and 16-byte loads / stores.
#pragma once
#include "mhdp_device.h"

namespace {

struct Row4 {
    int x0, y, z;
    uint32_t i0;
    uint32_t ym, zm;        // start of the rows y-1 and z-1 (wrapped; the row itself if none)
    uint32_t yp, zp;        // start of the rows y+1 and z+1 (wrapped; the row itself if none)
    uint32_t xm, xp;        // cells x0-1 and x0+4 (wrapped; x0 / x0+3 if none)
    bool ym_ok, zm_ok, yp_ok, zp_ok, xm_ok, xp_ok;
};
__device__ __forceinline__ Row4 row4(const Grid3& g, uint32_t t) {
    Row4 r;
    r.i0 = t << 2;
    int x, y, z;
    decode(g, r.i0, x, y, z);
    r.x0 = x;
    r.y = y;
    r.z = z;
    const uint32_t sy = (uint32_t)g.nx, sz = (uint32_t)g.nx * (uint32_t)g.ny;
    r.ym_ok = y > 0 || (g.per & 2u);
    r.ym = y > 0 ? r.i0 - sy : ((g.per & 2u) ? r.i0 + (uint32_t)(g.ny - 1) * sy : r.i0);
    r.yp_ok = y + 1 < g.ny || (g.per & 2u);
    r.yp = y + 1 < g.ny ? r.i0 + sy : ((g.per & 2u) ? r.i0 - (uint32_t)(g.ny - 1) * sy : r.i0);
    r.zm_ok = z > 0 || (g.per & 4u);
    r.zm = z > 0 ? r.i0 - sz : ((g.per & 4u) ? r.i0 + (uint32_t)(g.nz - 1) * sz : r.i0);
    r.zp_ok = z + 1 < g.nz || (g.per & 4u);
    r.zp = z + 1 < g.nz ? r.i0 + sz : ((g.per & 4u) ? r.i0 - (uint32_t)(g.nz - 1) * sz : r.i0);
    r.xm_ok = x > 0 || (g.per & 1u);
    r.xm = x > 0 ? r.i0 - 1 : ((g.per & 1u) ? r.i0 + (uint32_t)(g.nx - 1) : r.i0);
    r.xp_ok = x + 4 < g.nx || (g.per & 1u);
    r.xp = x + 4 < g.nx ? r.i0 + 4 : ((g.per & 1u) ? r.i0 + 4 - (uint32_t)g.nx : r.i0 + 3);
    return r;
}
__device__ __forceinline__ void ld4f(const float* __restrict__ p, uint32_t i, float (&v)[4]) {
    const float4 a = *reinterpret_cast<const float4*>(p + i);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void ld4u(const uint32_t* __restrict__ p, uint32_t i, uint32_t (&v)[4]) {
    const uint4 a = *reinterpret_cast<const uint4*>(p + i);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void st4f(float* __restrict__ p, uint32_t i, const float (&v)[4]) {
    *reinterpret_cast<float4*>(p + i) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void st4h(__hip_bfloat16* __restrict__ p, uint32_t i, const float (&v)[4]) {
    uint32_t a = (uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[0])) |
                 ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[1])) << 16);
    uint32_t b = (uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2])) |
                 ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[3])) << 16);
    *reinterpret_cast<uint2*>(p + i) = make_uint2(a, b);
}
__device__ __forceinline__ void ld4d(const double* __restrict__ p, uint32_t i, double (&v)[4]) {
    const double2 a = *reinterpret_cast<const double2*>(p + i), b = *reinterpret_cast<const double2*>(p + i + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void st4d(double* __restrict__ p, uint32_t i, const double (&v)[4]) {
    *reinterpret_cast<double2*>(p + i) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + i + 2) = make_double2(v[2], v[3]);
}

}  // namespace
