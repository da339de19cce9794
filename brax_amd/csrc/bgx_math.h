// Scalar, vector and quaternion helpers for the generalized-coordinates step.
//
// Each helper restates the reference's formula (brax/math.py, brax/base.py)
// with the same operation structure, so that results agree with the
// reference to rounding.  The vector types are templates on the scalar: the
// step runs on float; the adjoint's tangent pass runs the same formulas on
// Dual (value + directional derivative).
#pragma once

#ifdef BGX_HOST_CHECK
// A host build of the same device code for memory-safety tests on the CPU
// (tests/host/adj_tape_check.cpp, compiled by g++ with address and
// undefined-behaviour sanitizers); the library never defines this.
#include <cmath>
#define BGX_D inline
#else
#include <hip/hip_runtime.h>
#define BGX_D __device__ __forceinline__
#endif

namespace bgx {

// hardware reciprocal / square root / reciprocal square root (1 ulp); the
// step's divisions are all normalizations or scale factors whose rounding is
// far below the parity tolerance
#ifdef BGX_HOST_CHECK
BGX_D float rcp(float x) { return 1.f / x; }
BGX_D float fsqrt(float x) { return std::sqrt(x); }
BGX_D float rsq(float x) { return 1.f / std::sqrt(x); }
#else
BGX_D float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
BGX_D float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
BGX_D float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
#endif

// ---------------------------------------------------------------------------
// forward-mode dual number: v + d eps, eps^2 = 0

struct Dual {
  float v, d;
};
BGX_D Dual dual(float v, float d = 0.f) { return Dual{v, d}; }
BGX_D float val(float x) { return x; }
BGX_D float val(Dual x) { return x.v; }
BGX_D float tan_of(float) { return 0.f; }
BGX_D float tan_of(Dual x) { return x.d; }
template <class S>
BGX_D S lift(float x);
template <>
BGX_D float lift<float>(float x) { return x; }
template <>
BGX_D Dual lift<Dual>(float x) { return Dual{x, 0.f}; }

BGX_D Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
BGX_D Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
BGX_D Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
BGX_D Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
BGX_D Dual operator+(Dual a, float b) { return Dual{a.v + b, a.d}; }
BGX_D Dual operator+(float a, Dual b) { return Dual{a + b.v, b.d}; }
BGX_D Dual operator-(Dual a, float b) { return Dual{a.v - b, a.d}; }
BGX_D Dual operator-(float a, Dual b) { return Dual{a - b.v, -b.d}; }
BGX_D Dual operator*(Dual a, float b) { return Dual{a.v * b, a.d * b}; }
BGX_D Dual operator*(float a, Dual b) { return Dual{a * b.v, a * b.d}; }
BGX_D Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
BGX_D Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
BGX_D Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
// a zero tangent stays zero where the derivative is infinite (x = 0)
BGX_D Dual rcp(Dual x) {
  const float r = rcp(x.v);
  return Dual{r, x.d == 0.f ? 0.f : -x.d * (r * r)};
}
BGX_D Dual fsqrt(Dual x) {
  const float s = fsqrt(x.v);
  return Dual{s, x.d == 0.f ? 0.f : x.d * (0.5f * rcp(s))};
}
BGX_D Dual rsq(Dual x) {
  const float r = rsq(x.v);
  return Dual{r, x.d == 0.f ? 0.f : x.d * (-0.5f * r * r * r)};
}
BGX_D void sincos_(float x, float* s, float* c) { sincosf(x, s, c); }
BGX_D void sincos_(Dual x, Dual* s, Dual* c) {
  float sv, cv;
  sincosf(x.v, &sv, &cv);
  *s = Dual{sv, cv * x.d};
  *c = Dual{cv, -sv * x.d};
}
BGX_D float fabs_(float x) { return fabsf(x); }
BGX_D Dual fabs_(Dual x) { return x.v < 0.f ? -x : x; }
BGX_D float pow_(float x, float e) { return powf(x, e); }
BGX_D Dual pow_(Dual x, float e) {
  const float p = powf(x.v, e);
  return Dual{p, x.d * e * powf(x.v, e - 1.f)};
}
// min / max select on the value (the derivative follows the selected operand)
BGX_D float min_(float a, float b) { return fminf(a, b); }
BGX_D Dual min_(Dual a, Dual b) { return b.v < a.v ? b : a; }

// ---------------------------------------------------------------------------
// vectors and quaternions

template <class S>
struct V3T {
  S x, y, z;
};
template <class S>
struct Q4T {
  S w, x, y, z;
};
using V3 = V3T<float>;
using Q4 = Q4T<float>;

BGX_D V3 v3(float x, float y, float z) { return V3{x, y, z}; }
BGX_D V3 v3zero() { return V3{0.f, 0.f, 0.f}; }
BGX_D Q4 qident() { return Q4{1.f, 0.f, 0.f, 0.f}; }
template <class S>
BGX_D V3T<S> lift3(V3 a) {
  return V3T<S>{lift<S>(a.x), lift<S>(a.y), lift<S>(a.z)};
}
template <class S>
BGX_D Q4T<S> lift4(Q4 a) {
  return Q4T<S>{lift<S>(a.w), lift<S>(a.x), lift<S>(a.y), lift<S>(a.z)};
}
template <class S>
BGX_D V3 val3(V3T<S> a) {
  return V3{val(a.x), val(a.y), val(a.z)};
}

template <class S>
BGX_D V3T<S> operator+(V3T<S> a, V3T<S> b) {
  return V3T<S>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class S>
BGX_D V3T<S> operator-(V3T<S> a, V3T<S> b) {
  return V3T<S>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class S>
BGX_D V3T<S> operator-(V3T<S> a) {
  return V3T<S>{-a.x, -a.y, -a.z};
}
template <class S>
BGX_D V3T<S> operator*(S s, V3T<S> a) {
  return V3T<S>{s * a.x, s * a.y, s * a.z};
}
template <class S>
BGX_D V3T<S> operator*(V3T<S> a, S s) {
  return V3T<S>{a.x * s, a.y * s, a.z * s};
}
// a dual vector scaled by a float, and a float vector scaled by a dual
BGX_D V3T<Dual> operator*(V3T<Dual> a, float s) {
  return V3T<Dual>{a.x * s, a.y * s, a.z * s};
}
BGX_D V3T<Dual> operator*(float s, V3T<Dual> a) {
  return V3T<Dual>{s * a.x, s * a.y, s * a.z};
}
BGX_D V3T<Dual> operator*(V3 a, Dual s) {
  return V3T<Dual>{a.x * s, a.y * s, a.z * s};
}
template <class S>
BGX_D V3T<S>& operator+=(V3T<S>& a, V3T<S> b) {
  a = a + b;
  return a;
}
template <class S>
BGX_D V3T<S>& operator-=(V3T<S>& a, V3T<S> b) {
  a = a - b;
  return a;
}
template <class S>
BGX_D S dot(V3T<S> a, V3T<S> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class S>
BGX_D V3T<S> cross(V3T<S> a, V3T<S> b) {
  return V3T<S>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
                a.x * b.y - a.y * b.x};
}
template <class S>
BGX_D V3T<S> qvec(Q4T<S> q) {
  return V3T<S>{q.x, q.y, q.z};
}

// math.py:25-40 rotate(vec, quat)
template <class S>
BGX_D V3T<S> rotate(V3T<S> v, Q4T<S> q) {
  const V3T<S> u = qvec(q);
  const S s = q.w;
  V3T<S> r = ((2.f * dot(u, v)) * u) + ((s * s - dot(u, u)) * v);
  return r + ((2.f * s) * cross(u, v));
}

// math.py:86-101 quat_mul
template <class S>
BGX_D Q4T<S> qmul(Q4T<S> u, Q4T<S> v) {
  return Q4T<S>{u.w * v.w - u.x * v.x - u.y * v.y - u.z * v.z,
                u.w * v.x + u.x * v.w + u.y * v.z - u.z * v.y,
                u.w * v.y - u.x * v.z + u.y * v.w + u.z * v.x,
                u.w * v.z + u.x * v.y - u.y * v.x + u.z * v.w};
}

// math.py:122-131 quat_inv
template <class S>
BGX_D Q4T<S> qinv(Q4T<S> q) {
  return Q4T<S>{q.w, -q.x, -q.y, -q.z};
}

// math.py:134-148 quat_rot_axis
template <class S>
BGX_D Q4T<S> qrot_axis(V3T<S> axis, S angle) {
  S s, c;
  sincos_(angle * 0.5f, &s, &c);
  return Q4T<S>{c, axis.x * s, axis.y * s, axis.z * s};
}

// math.py:308-345 safe_norm + normalize: x / (|x| + 1e-6 [|x| == 0]); the
// norm is 0 when every component is within 1e-8 of zero (jp.allclose).
template <class S>
BGX_D Q4T<S> qnormalize(Q4T<S> q) {
  const bool zero = fabsf(val(q.w)) <= 1e-8f && fabsf(val(q.x)) <= 1e-8f &&
                    fabsf(val(q.y)) <= 1e-8f && fabsf(val(q.z)) <= 1e-8f;
  const S n =
      zero ? lift<S>(0.f) : fsqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  const S r = rcp(n + (zero ? 1e-6f : 0.f));
  return Q4T<S>{q.w * r, q.x * r, q.y * r, q.z * r};
}

template <class S>
BGX_D V3T<S> vnormalize(V3T<S> v, S* norm_out) {
  const bool zero = fabsf(val(v.x)) <= 1e-8f && fabsf(val(v.y)) <= 1e-8f &&
                    fabsf(val(v.z)) <= 1e-8f;
  const S n = zero ? lift<S>(0.f) : fsqrt(dot(v, v));
  const S r = rcp(n + (zero ? 1e-6f : 0.f));
  if (norm_out) *norm_out = n;
  return V3T<S>{v.x * r, v.y * r, v.z * r};
}

// math.py:151-165 quat_to_3x3 (m[row][col])
template <class S>
BGX_D void quat_to_3x3(Q4T<S> q, S m[3][3]) {
  const S d = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  const S s = 2.f * rcp(d);
  const S xs = q.x * s, ys = q.y * s, zs = q.z * s;
  const S wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
  const S xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
  const S yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
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
template <class S>
BGX_D V3T<S> quat_zaxis(Q4T<S> q) {
  S m[3][3];
  quat_to_3x3(q, m);
  return V3T<S>{m[0][2], m[1][2], m[2][2]};
}

// base.py:604-608 Motion.cross(Motion): returns (ang, vel)
template <class S>
BGX_D void cross_motion(V3T<S> sa, V3T<S> sv, V3T<S> ma, V3T<S> mv, V3T<S>& oa,
                        V3T<S>& ov) {
  oa = cross(sa, ma);
  ov = cross(sa, mv) + cross(sv, ma);
}

// base.py:611-614 Motion.cross(Force): returns (ang, vel)
template <class S>
BGX_D void cross_force(V3T<S> sa, V3T<S> sv, V3T<S> fa, V3T<S> fv, V3T<S>& oa,
                       V3T<S>& ov) {
  oa = cross(sa, fa) + cross(sv, fv);
  ov = cross(sa, fv);
}

// symmetric 3x3 stored as xx yy zz xy xz yz
template <class S>
struct Sym3T {
  S xx, yy, zz, xy, xz, yz;
};
using Sym3 = Sym3T<float>;
template <class S>
BGX_D V3T<S> symmul(const Sym3T<S>& i, V3T<S> a) {
  return V3T<S>{i.xx * a.x + i.xy * a.y + i.xz * a.z,
                i.xy * a.x + i.yy * a.y + i.yz * a.z,
                i.xz * a.x + i.yz * a.y + i.zz * a.z};
}
template <class S>
BGX_D Sym3T<S> symadd(const Sym3T<S>& a, const Sym3T<S>& b) {
  return Sym3T<S>{a.xx + b.xx, a.yy + b.yy, a.zz + b.zz,
                  a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}

// base.py:297-302 Inertia.mul(Motion) with c = transform.pos (mass-weighted
// COM offset): returns the force (ang, vel)
template <class S, class MS>
BGX_D void inertia_mul(V3T<S> c, const Sym3T<S>& i, MS mass, V3T<S> ma,
                       V3T<S> mv, V3T<S>& fa, V3T<S>& fv) {
  fa = symmul(i, ma) + cross(c, mv);
  fv = (mv * mass) - cross(c, ma);
}

}  // namespace bgx
