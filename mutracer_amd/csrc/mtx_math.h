// Device/host f32 vector math for the mtx kernels (gfx950).
//
// Arithmetic contract (the whole library is compiled with -ffp-contract=off,
// correctly rounded division and square root, denormals kept):
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)   = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   normalize(v) = v / sqrt(dot(v,v))
//   M*v          = ((c0*v.x + c1*v.y) + c2*v.z) + c3*v.w   (column-major 4x4)
//   min/max      = minNum/maxNum, first operand kept when equal
// WGSL leaves these orderings open; fixing them makes the traversal's
// closest-hit IDs reproducible bit-for-bit on any IEEE f32 implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MTX_HD __host__ __device__ __forceinline__

namespace mtx {

constexpr float kPi = 3.14159265359f;
constexpr float kTwoPi = 6.28318530718f;
constexpr float kInvPi = 0.31830988618f;
constexpr float kInvPi2 = 0.10132118364f;
constexpr float kEps = 0.0005f;
constexpr float kInf = 1.0e+38f;
constexpr float kRoughnessMin = 0.001f;
constexpr float kRayMaxDistance = 9999.0f;

struct f3 {
  float x, y, z;
};
MTX_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
MTX_HD f3 splat(float v) { return f3{v, v, v}; }
MTX_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MTX_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MTX_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
MTX_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
MTX_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
MTX_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
MTX_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
MTX_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
MTX_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
MTX_HD f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
MTX_HD float len(f3 v) { return sqrtf(dot(v, v)); }
MTX_HD f3 normalize(f3 v) { return v / len(v); }
MTX_HD float mnf(float a, float b) { return (b < a || a != a) ? b : a; }
MTX_HD float mxf(float a, float b) { return (b > a || a != a) ? b : a; }
MTX_HD f3 vmin(f3 a, f3 b) { return f3{mnf(a.x, b.x), mnf(a.y, b.y), mnf(a.z, b.z)}; }
MTX_HD f3 vmax(f3 a, f3 b) { return f3{mxf(a.x, b.x), mxf(a.y, b.y), mxf(a.z, b.z)}; }
MTX_HD f3 vabs(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
MTX_HD float clampf(float v, float lo, float hi) { return mnf(mxf(v, lo), hi); }
MTX_HD float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
MTX_HD f3 mix3(f3 a, f3 b, float t) { return a * (1.0f - t) + b * t; }
MTX_HD f3 mix3v(f3 a, f3 b, f3 t) { return a * (splat(1.0f) - t) + b * t; }
MTX_HD float maxc(f3 v) { return mxf(mxf(v.x, v.y), v.z); }
MTX_HD f3 sel(bool c, f3 a, f3 b) { return c ? a : b; }
MTX_HD float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
MTX_HD f3 f3of(float4 v) { return f3{v.x, v.y, v.z}; }

// Column-major affine transform held as four xyz columns.
struct Affine {
  f3 c0, c1, c2, c3;
};
MTX_HD f3 xform(const Affine &m, f3 v, float w) {
  f3 r;
  r.x = ((m.c0.x * v.x + m.c1.x * v.y) + m.c2.x * v.z) + m.c3.x * w;
  r.y = ((m.c0.y * v.x + m.c1.y * v.y) + m.c2.y * v.z) + m.c3.y * w;
  r.z = ((m.c0.z * v.x + m.c1.z * v.y) + m.c2.z * v.z) + m.c3.z * w;
  return r;
}
// transpose(mat3(c0,c1,c2)) * n
MTX_HD f3 xform_normal_transposed(const Affine &m, f3 n) {
  return f3{dot(m.c0, n), dot(m.c1, n), dot(m.c2, n)};
}

MTX_HD uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
MTX_HD float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }
MTX_HD int32_t fint(float f) { return __builtin_bit_cast(int32_t, f); }
MTX_HD float intf(int32_t u) { return __builtin_bit_cast(float, u); }

// Order-preserving float <-> uint key for atomic min/max (-0 sorts below +0).
MTX_HD uint32_t fkey(float f) {
  uint32_t u = fbits(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
MTX_HD float unkey(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return bitsf(u);
}

}  // namespace mtx
