// ORACLE — test infrastructure only. Nothing in the product library links,
// includes or calls anything under oracle/; only tests/, the smoke check and
// bench.py's CPU-baseline leg load it, as the checker.
//
// Three-component f32 vector with the arithmetic of the reference's `glam::Vec3`
// (glam 0.33.6, Cargo.lock:101-105): every operation is the plain scalar IEEE
// f32 expression, evaluated left to right, never contracted to FMA (the
// reference is built for baseline x86-64 with no FMA).  Build with
// -ffp-contract=off and without -ffast-math.
//
// Formulas assumed for glam (its source is not in the reference tree):
//   dot        (x*x' + y*y') + z*z'
//   length     sqrt(dot(v,v))
//   normalize  v * (1 / length(v))
//   v / s      per-component division
//   cross      (y*z' - y'*z, z*x' - z'*x, x*y' - x'*y)
//   lerp       a*(1-s) + b*s
//   max/min    per component, NaN-ignoring (Rust f32::max/min == fmaxf/fminf)
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

struct V3 {
    float x, y, z;
};

static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 splat(float s) { return V3{s, s, s}; }
static inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
static inline V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
static inline V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
static inline V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
static inline V3& operator*=(V3& a, float s) { a = a * s; return a; }
static inline V3& operator/=(V3& a, float s) { a = a / s; return a; }

static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float length_sq(V3 a) { return dot(a, a); }
static inline float length(V3 a) { return std::sqrt(dot(a, a)); }
static inline V3 normalize(V3 a) { float r = 1.0f / length(a); return a * r; }
static inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// Rust f32::max / f32::min: if one operand is NaN the other is returned.
static inline float fmax_(float a, float b) { return std::fmax(a, b); }
static inline float fmin_(float a, float b) { return std::fmin(a, b); }
// Rust f32::clamp: NaN passes through.
static inline float clampf(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
static inline V3 vabs(V3 a) { return V3{std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
static inline V3 vmax(V3 a, V3 b) { return V3{fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
static inline V3 vmin(V3 a, V3 b) { return V3{fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
static inline float max_elem(V3 a) { return fmax_(a.x, fmax_(a.y, a.z)); }
static inline float min_elem(V3 a) { return fmin_(a.x, fmin_(a.y, a.z)); }
static inline V3 lerp(V3 a, V3 b, float s) { return a * (1.0f - s) + b * s; }
static inline float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Rust `as` casts from f32 saturate and map NaN to 0.
static inline int32_t as_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}
static inline uint32_t as_u32(float f) {
    if (!(f > 0.0f)) return 0;  // NaN, zero and negatives
    if (f >= 4294967296.0f) return UINT32_MAX;
    return (uint32_t)f;
}
static inline uint64_t as_u64(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)f;
}
static inline uint8_t as_u8(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 255.0f) return 255;
    return (uint8_t)f;
}
static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

static const float PI_F = 3.14159265358979323846f;
static const float TAU_F = 6.28318530717958647692f;

}  // namespace orc
