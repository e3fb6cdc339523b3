// CPU oracle for the mutracer render hot path -- TEST INFRASTRUCTURE ONLY.
//
// This directory restates the reference algorithms on the CPU so that the HIP
// path can be checked against them. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker or as the
// reported CPU baseline. Nothing in mutracer_amd/ includes, links or calls it.
//
// Arithmetic contract shared with the HIP kernels (both sides are compiled
// with -ffp-contract=off, IEEE f32, correctly rounded / and sqrt):
//   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)    = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize(v)  = v / sqrt(dot(v,v))           (component-wise division)
//   M*v (4x4, column-major, WGSL mat4x4*vec4) = ((c0*v.x + c1*v.y) + c2*v.z) + c3*v.w
//   min/max       = minNum/maxNum, first operand kept when equal (mnf/mxf)
// WGSL leaves these orderings to the implementation; the oracle fixes them.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

constexpr float PI = 3.14159265359f;
constexpr float TWOPI = 6.28318530718f;
constexpr float INVPI = 0.31830988618f;
constexpr float INVPI2 = 0.10132118364f;
constexpr float EPS = 0.0005f;                 // common.wgsl.ts:7
constexpr float INF = 1.0e+38f;                // common.wgsl.ts:9
constexpr float ROUGHNESS_MIN = 0.001f;        // common.wgsl.ts:10
constexpr float RAY_MAX_DISTANCE = 9999.0f;    // rayTraceCore.wgsl.ts:46

struct V3 {
  float x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(float a, float b, float c) : x(a), y(b), z(c) {}
  explicit V3(float a) : x(a), y(a), z(a) {}
  float operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
  float &operator[](int i) { return i == 0 ? x : i == 1 ? y : z; }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, V3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(V3 v) { return std::sqrt(dot(v, v)); }
inline V3 normalize(V3 v) { return v / length(v); }
// minNum/maxNum with a fixed rule for equal operands (first operand wins) so
// that the HIP side can use the identical select sequence.
inline float mnf(float a, float b) { return (b < a || a != a) ? b : a; }
inline float mxf(float a, float b) { return (b > a || a != a) ? b : a; }
inline V3 vmin(V3 a, V3 b) { return V3(mnf(a.x, b.x), mnf(a.y, b.y), mnf(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return V3(mxf(a.x, b.x), mxf(a.y, b.y), mxf(a.z, b.z)); }
inline V3 vabs(V3 a) { return V3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
inline float clampf(float v, float lo, float hi) { return mnf(mxf(v, lo), hi); }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline V3 mix3(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
inline V3 mix3(V3 a, V3 b, V3 t) { return a * (V3(1.0f) - t) + b * t; }
inline float maxc(V3 v) { return mxf(mxf(v.x, v.y), v.z); }
inline V3 sel(bool c, V3 a, V3 b) { return c ? a : b; }

// WGSL mat4x4<f32> * vec4<f32> for a column-major 16-float matrix, xyz only.
inline V3 mat_mul_point(const float *m, V3 p, float w) {
  V3 r;
  for (int k = 0; k < 3; k++) {
    float s = m[0 + k] * p.x;
    s = s + m[4 + k] * p.y;
    s = s + m[8 + k] * p.z;
    s = s + m[12 + k] * w;
    r[k] = s;
  }
  return r;
}
// Full vec4 result of M*v.
inline void mat_mul4(const float *m, const float v[4], float out[4]) {
  for (int k = 0; k < 4; k++) {
    float s = m[0 + k] * v[0];
    s = s + m[4 + k] * v[1];
    s = s + m[8 + k] * v[2];
    s = s + m[12 + k] * v[3];
    out[k] = s;
  }
}

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline int32_t f2i(float f) { int32_t u; std::memcpy(&u, &f, 4); return u; }
inline float i2f(int32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

}  // namespace orc
