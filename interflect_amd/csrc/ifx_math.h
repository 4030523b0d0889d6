// Scalar f32 vector arithmetic shared by the gfx950 kernels and the host-side
// surfel generator.  Every expression keeps the operand order of the
// reference's glam::Vec3 code (src/*.rs) and is compiled with
// -ffp-contract=off and correctly rounded f32 division / sqrt, so a value
// computed here is bit-identical to the same expression in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IFX_HD __host__ __device__ __forceinline__

namespace ifx {

struct f3 {
    float x, y, z;
};

IFX_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
IFX_HD f3 spl(float s) { return f3{s, s, s}; }
IFX_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
IFX_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
IFX_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
IFX_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
IFX_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
IFX_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
IFX_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
IFX_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
IFX_HD float len3(f3 a) { return sqrtf(dot3(a, a)); }
IFX_HD f3 nrm3(f3 a) {
    float r = 1.0f / len3(a);
    return a * r;
}
IFX_HD f3 cross3(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
IFX_HD float fmx(float a, float b) { return fmaxf(a, b); }
IFX_HD float fmn(float a, float b) { return fminf(a, b); }
// Rust f32::clamp semantics (NaN passes through).
IFX_HD float clmp(float x, float lo, float hi) {
    x = x < lo ? lo : x;
    return x > hi ? hi : x;
}
IFX_HD f3 abs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
IFX_HD f3 max3(f3 a, f3 b) { return f3{fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }
IFX_HD f3 min3(f3 a, f3 b) { return f3{fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
IFX_HD float maxe(f3 a) { return fmx(a.x, fmx(a.y, a.z)); }
IFX_HD float mine(f3 a) { return fmn(a.x, fmn(a.y, a.z)); }
IFX_HD f3 lerp3(f3 a, f3 b, float s) { return a * (1.0f - s) + b * s; }
IFX_HD float at3(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
IFX_HD f3 ld3(float4 v) { return f3{v.x, v.y, v.z}; }
IFX_HD float4 st4(f3 v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }

// Saturating float -> int conversions with NaN -> 0 (Rust `as`).
IFX_HD int32_t sat_i32(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)f;
}
IFX_HD uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}
IFX_HD int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;

}  // namespace ifx
