// Vector and quaternion helpers for the generalized-coordinates step.
//
// Each helper restates the reference's formula (brax/math.py, brax/base.py)
// in FP32 with the same operation structure, so that results agree with the
// reference to rounding.
#pragma once

#include <hip/hip_runtime.h>

#define BGX_D __device__ __forceinline__

namespace bgx {

struct V3 {
  float x, y, z;
};
struct Q4 {
  float w, x, y, z;
};

// hardware reciprocal / square root / reciprocal square root (1 ulp); the
// step's divisions are all normalizations or scale factors whose rounding is
// far below the parity tolerance
BGX_D float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
BGX_D float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
BGX_D float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

BGX_D V3 v3(float x, float y, float z) { return V3{x, y, z}; }
BGX_D V3 v3zero() { return V3{0.f, 0.f, 0.f}; }
BGX_D Q4 qident() { return Q4{1.f, 0.f, 0.f, 0.f}; }
BGX_D V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BGX_D V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
BGX_D V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
BGX_D V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
BGX_D V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
BGX_D V3& operator+=(V3& a, V3 b) {
  a = a + b;
  return a;
}
BGX_D V3& operator-=(V3& a, V3 b) {
  a = a - b;
  return a;
}
BGX_D float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BGX_D V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
BGX_D V3 qvec(Q4 q) { return V3{q.x, q.y, q.z}; }

// math.py:25-40 rotate(vec, quat)
BGX_D V3 rotate(V3 v, Q4 q) {
  const V3 u = qvec(q);
  const float s = q.w;
  V3 r = (2.f * dot(u, v)) * u + (s * s - dot(u, u)) * v;
  return r + (2.f * s) * cross(u, v);
}

// math.py:86-101 quat_mul
BGX_D Q4 qmul(Q4 u, Q4 v) {
  return Q4{u.w * v.w - u.x * v.x - u.y * v.y - u.z * v.z,
            u.w * v.x + u.x * v.w + u.y * v.z - u.z * v.y,
            u.w * v.y - u.x * v.z + u.y * v.w + u.z * v.x,
            u.w * v.z + u.x * v.y - u.y * v.x + u.z * v.w};
}

// math.py:122-131 quat_inv
BGX_D Q4 qinv(Q4 q) { return Q4{q.w, -q.x, -q.y, -q.z}; }

// math.py:134-148 quat_rot_axis
BGX_D Q4 qrot_axis(V3 axis, float angle) {
  float s, c;
  sincosf(angle * 0.5f, &s, &c);
  return Q4{c, axis.x * s, axis.y * s, axis.z * s};
}

// math.py:308-345 safe_norm + normalize: x / (|x| + 1e-6 [|x| == 0]); the
// norm is 0 when every component is within 1e-8 of zero (jp.allclose).
BGX_D Q4 qnormalize(Q4 q) {
  const bool zero = fabsf(q.w) <= 1e-8f && fabsf(q.x) <= 1e-8f &&
                    fabsf(q.y) <= 1e-8f && fabsf(q.z) <= 1e-8f;
  const float n =
      zero ? 0.f : fsqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  const float r = rcp(n + (n == 0.f ? 1e-6f : 0.f));
  return Q4{q.w * r, q.x * r, q.y * r, q.z * r};
}

BGX_D V3 vnormalize(V3 v, float* norm_out) {
  const bool zero = fabsf(v.x) <= 1e-8f && fabsf(v.y) <= 1e-8f && fabsf(v.z) <= 1e-8f;
  const float n = zero ? 0.f : fsqrt(dot(v, v));
  const float r = rcp(n + (n == 0.f ? 1e-6f : 0.f));
  if (norm_out) *norm_out = n;
  return V3{v.x * r, v.y * r, v.z * r};
}

// math.py:151-165 quat_to_3x3 (m[row][col])
BGX_D void quat_to_3x3(Q4 q, float m[3][3]) {
  const float d = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  const float s = 2.f * rcp(d);
  const float xs = q.x * s, ys = q.y * s, zs = q.z * s;
  const float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
  const float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
  const float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
  m[0][0] = 1.f - (yy + zz);
  m[0][1] = xy - wz;
  m[0][2] = xz + wy;
  m[1][0] = xy + wz;
  m[1][1] = 1.f - (xx + zz);
  m[1][2] = yz - wx;
  m[2][0] = xz - wy;
  m[2][1] = yz + wx;
  m[2][2] = 1.f - (xx + yy);
}

// third column of quat_to_3x3 (the frame's z axis)
BGX_D V3 quat_zaxis(Q4 q) {
  float m[3][3];
  quat_to_3x3(q, m);
  return V3{m[0][2], m[1][2], m[2][2]};
}

// base.py:604-608 Motion.cross(Motion): returns (ang, vel)
BGX_D void cross_motion(V3 sa, V3 sv, V3 ma, V3 mv, V3& oa, V3& ov) {
  oa = cross(sa, ma);
  ov = cross(sa, mv) + cross(sv, ma);
}

// base.py:611-614 Motion.cross(Force): returns (ang, vel)
BGX_D void cross_force(V3 sa, V3 sv, V3 fa, V3 fv, V3& oa, V3& ov) {
  oa = cross(sa, fa) + cross(sv, fv);
  ov = cross(sa, fv);
}

// symmetric 3x3 stored as xx yy zz xy xz yz
struct Sym3 {
  float xx, yy, zz, xy, xz, yz;
};
BGX_D V3 symmul(const Sym3& i, V3 a) {
  return V3{i.xx * a.x + i.xy * a.y + i.xz * a.z,
            i.xy * a.x + i.yy * a.y + i.yz * a.z,
            i.xz * a.x + i.yz * a.y + i.zz * a.z};
}
BGX_D Sym3 symadd(const Sym3& a, const Sym3& b) {
  return Sym3{a.xx + b.xx, a.yy + b.yy, a.zz + b.zz,
              a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}

// base.py:297-302 Inertia.mul(Motion) with c = transform.pos (mass-weighted
// COM offset): returns the force (ang, vel)
BGX_D void inertia_mul(V3 c, const Sym3& i, float mass, V3 ma, V3 mv, V3& fa,
                       V3& fv) {
  fa = symmul(i, ma) + cross(c, mv);
  fv = mass * mv - cross(c, ma);
}

}  // namespace bgx
