// Register-level helpers of the fine-grid stencil passes: four-wide float views of FP32 and
// BF16 rows, and raw loads that are issued up front and pinned in registers (see R4 below).
#pragma once
#include "mhdp_device.h"

namespace {

struct F4 {
    float v[4];
};
__device__ __forceinline__ float bfl(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bfh(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }
__device__ __forceinline__ F4 ld4(const float* __restrict__ p, uint32_t i) {
    const float4 a = *reinterpret_cast<const float4*>(p + i);
    return F4{{a.x, a.y, a.z, a.w}};
}
__device__ __forceinline__ F4 ld4(const __hip_bfloat16* __restrict__ p, uint32_t i) {
    const uint2 a = *reinterpret_cast<const uint2*>(p + i);
    return F4{{bfl(a.x), bfh(a.x), bfl(a.y), bfh(a.y)}};
}
__device__ __forceinline__ void ld2(const float* __restrict__ p, uint32_t i, float& a, float& b) {
    const float2 v = *reinterpret_cast<const float2*>(p + i);
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void ld2(const __hip_bfloat16* __restrict__ p, uint32_t i, float& a, float& b) {
    const uint32_t v = *reinterpret_cast<const uint32_t*>(p + i);
    a = bfl(v);
    b = bfh(v);
}
__device__ __forceinline__ void st4(float* __restrict__ p, uint32_t i, const F4& a) {
    *reinterpret_cast<float4*>(p + i) = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]);
}
__device__ __forceinline__ uint32_t pk(float a, float b) {
    return (uint32_t)__bfloat16_as_ushort(__float2bfloat16(a)) |
           ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(b)) << 16);
}
__device__ __forceinline__ void st4(__hip_bfloat16* __restrict__ p, uint32_t i, const F4& a) {
    *reinterpret_cast<uint2*>(p + i) = make_uint2(pk(a.v[0], a.v[1]), pk(a.v[2], a.v[3]));
}
__device__ __forceinline__ void st2(float* __restrict__ p, uint32_t i, float a, float b) {
    *reinterpret_cast<float2*>(p + i) = make_float2(a, b);
}
__device__ __forceinline__ void st2(__hip_bfloat16* __restrict__ p, uint32_t i, float a, float b) {
    *reinterpret_cast<uint32_t*>(p + i) = pk(a, b);
}

// ---- loads issued up front.  Every value a brick needs is loaded unconditionally (all
// indices are in range by construction) into raw registers, then each raw register passes an
// empty asm statement that claims to modify it: the compiler can then neither sink a load
// into a data-dependent branch nor wait for the loads one group at a time, so one brick costs
// one memory round trip.
template <class T> struct R4;
template <> struct R4<float> { float4 v; };
template <> struct R4<__hip_bfloat16> { uint2 v; };
template <class T> struct R1;
template <> struct R1<float> { float v; };
template <> struct R1<__hip_bfloat16> { uint32_t v; };
template <class T> struct R2;
template <> struct R2<float> { float2 v; };
template <> struct R2<__hip_bfloat16> { uint32_t v; };

__device__ __forceinline__ R4<float> ldr4(const float* __restrict__ p, uint32_t i) {
    return R4<float>{*reinterpret_cast<const float4*>(p + i)};
}
__device__ __forceinline__ R4<__hip_bfloat16> ldr4(const __hip_bfloat16* __restrict__ p, uint32_t i) {
    return R4<__hip_bfloat16>{*reinterpret_cast<const uint2*>(p + i)};
}
__device__ __forceinline__ R1<float> ldr1(const float* __restrict__ p, uint32_t i) { return R1<float>{p[i]}; }
__device__ __forceinline__ R1<__hip_bfloat16> ldr1(const __hip_bfloat16* __restrict__ p, uint32_t i) {
    return R1<__hip_bfloat16>{(uint32_t)*reinterpret_cast<const unsigned short*>(p + i)};
}
__device__ __forceinline__ R2<float> ldr2(const float* __restrict__ p, uint32_t i) {
    return R2<float>{*reinterpret_cast<const float2*>(p + i)};
}
__device__ __forceinline__ R2<__hip_bfloat16> ldr2(const __hip_bfloat16* __restrict__ p, uint32_t i) {
    return R2<__hip_bfloat16>{*reinterpret_cast<const uint32_t*>(p + i)};
}
__device__ __forceinline__ void fence(R4<float>& r) {
    asm volatile("" : "+v"(r.v.x), "+v"(r.v.y), "+v"(r.v.z), "+v"(r.v.w));
}
__device__ __forceinline__ void fence(R4<__hip_bfloat16>& r) { asm volatile("" : "+v"(r.v.x), "+v"(r.v.y)); }
__device__ __forceinline__ void fence(R1<float>& r) { asm volatile("" : "+v"(r.v)); }
__device__ __forceinline__ void fence(R1<__hip_bfloat16>& r) { asm volatile("" : "+v"(r.v)); }
__device__ __forceinline__ void fence(R2<float>& r) { asm volatile("" : "+v"(r.v.x), "+v"(r.v.y)); }
__device__ __forceinline__ void fence(R2<__hip_bfloat16>& r) { asm volatile("" : "+v"(r.v)); }
__device__ __forceinline__ F4 un(const R4<float>& r) { return F4{{r.v.x, r.v.y, r.v.z, r.v.w}}; }
__device__ __forceinline__ F4 un(const R4<__hip_bfloat16>& r) {
    return F4{{bfl(r.v.x), bfh(r.v.x), bfl(r.v.y), bfh(r.v.y)}};
}
__device__ __forceinline__ float un(const R1<float>& r) { return r.v; }
__device__ __forceinline__ float un(const R1<__hip_bfloat16>& r) { return bfl(r.v); }
__device__ __forceinline__ void un(const R2<float>& r, float& a, float& b) { a = r.v.x; b = r.v.y; }
__device__ __forceinline__ void un(const R2<__hip_bfloat16>& r, float& a, float& b) { a = bfl(r.v); b = bfh(r.v); }

}  // namespace
