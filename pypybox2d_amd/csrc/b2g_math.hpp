// b2g_math.hpp -- 2-D vector math for the contact pipeline kernels.
//
// The operations follow pypybox2d/common.py (PyVec2, PyMat22, PyTransform,
// PyAABB) in meaning and in evaluation order:
//   Vec2 * scalar  -> (s*x, s*y)                    common.py:113-117
//   v.cross(s)     -> (s*y, -s*x)                   common.py:200-204
//   scalar_cross   -> (-s*y, s*x)                   common.py:188-195
//   Transform * v  -> (p.x + c1.x*v.x) + c2.x*v.y   common.py:677-685
//   min(a, b) / max(a, b) keep the first argument on ties.
// Everything is templated on the scalar type T (float for the FP32-bulk
// pipeline, double for the FP64 pipeline).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define B2G_HD __host__ __device__ __forceinline__

namespace b2g {

template <typename T>
struct V2 {
  T x, y;
};

template <typename T>
B2G_HD V2<T> v2(T x, T y) {
  V2<T> r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename T>
B2G_HD V2<T> operator+(V2<T> a, V2<T> b) { return v2<T>(a.x + b.x, a.y + b.y); }
template <typename T>
B2G_HD V2<T> operator-(V2<T> a, V2<T> b) { return v2<T>(a.x - b.x, a.y - b.y); }
template <typename T>
B2G_HD V2<T> operator-(V2<T> a) { return v2<T>(-a.x, -a.y); }
template <typename T>
B2G_HD V2<T> scale(T s, V2<T> a) { return v2<T>(s * a.x, s * a.y); }
template <typename T>
B2G_HD T dot(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
template <typename T>
B2G_HD T cross(V2<T> a, V2<T> b) { return a.x * b.y - a.y * b.x; }
template <typename T>
B2G_HD V2<T> cross_vs(V2<T> a, T s) { return v2<T>(s * a.y, -s * a.x); }
template <typename T>
B2G_HD V2<T> cross_sv(T s, V2<T> a) { return v2<T>(-s * a.y, s * a.x); }
template <typename T>
B2G_HD T sq(T x) { return x * x; }
template <typename T>
B2G_HD T pmin(T a, T b) { return (b < a) ? b : a; }
template <typename T>
B2G_HD T pmax(T a, T b) { return (b > a) ? b : a; }
template <typename T>
B2G_HD T pclamp(T v, T lo, T hi) { return pmax(lo, pmin(v, hi)); }
template <typename T>
B2G_HD V2<T> vmin(V2<T> a, V2<T> b) { return v2<T>(pmin(a.x, b.x), pmin(a.y, b.y)); }
template <typename T>
B2G_HD V2<T> vmax(V2<T> a, V2<T> b) { return v2<T>(pmax(a.x, b.x), pmax(a.y, b.y)); }

template <typename To, typename From>
B2G_HD V2<To> cvt(V2<From> a) { return v2<To>((To)a.x, (To)a.y); }

// Vec2.normalized (common.py:216-222): below eps the vector is returned as is.
template <typename T>
struct NormEps;
template <>
struct NormEps<double> {
  static constexpr double value = 2.220446049250313080847e-16;  // DBL_EPSILON
};
template <>
struct NormEps<float> {
  static constexpr float value = 1.1920928955078125e-07f;  // FLT_EPSILON
};

template <typename T>
B2G_HD V2<T> normalized(V2<T> a) {
  T l = sqrt(sq(a.x) + sq(a.y));
  if (l < NormEps<T>::value) return a;
  T il = T(1) / l;
  return v2<T>(a.x * il, a.y * il);
}

// Rotation R = [[c, -s], [s, c]] (PyMat22.angle setter, common.py:492-500).
template <typename T>
struct Rot {
  T c, s;
};
template <typename T>
B2G_HD V2<T> rmul(Rot<T> r, V2<T> v) {  // R * v
  return v2<T>(r.c * v.x + (-r.s) * v.y, r.s * v.x + r.c * v.y);
}
template <typename T>
B2G_HD V2<T> rmul_t(Rot<T> r, V2<T> v) {  // R^T * v = (col1.v, col2.v)
  return v2<T>(r.c * v.x + r.s * v.y, (-r.s) * v.x + r.c * v.y);
}

template <typename T>
struct XF {
  V2<T> p;
  Rot<T> q;
};
template <typename T>
B2G_HD V2<T> xmul(XF<T> t, V2<T> o) {  // Transform * v
  return v2<T>(t.p.x + t.q.c * o.x + (-t.q.s) * o.y, t.p.y + t.q.s * o.x + t.q.c * o.y);
}
template <typename T>
B2G_HD V2<T> xmul_t(XF<T> t, V2<T> v) {  // Transform.mul_t(v) = R^T (v - p)
  return rmul_t(t.q, v - t.p);
}

// 2x2 matrix, column major (c1 = (a11, a21), c2 = (a12, a22)).
template <typename T>
struct M22 {
  T c1x, c1y, c2x, c2y;
};
template <typename T>
B2G_HD V2<T> mmul(M22<T> m, V2<T> v) {
  return v2<T>(m.c1x * v.x + m.c2x * v.y, m.c1y * v.x + m.c2y * v.y);
}
template <typename T>
B2G_HD V2<T> mmul_t(M22<T> m, V2<T> v) {
  return v2<T>(m.c1x * v.x + m.c1y * v.y, m.c2x * v.x + m.c2y * v.y);
}
// Transform with a general rotation matrix (the edge/polygon collider builds
// one with Transform.mul_t(Transform), common.py:670-673).
template <typename T>
struct XFM {
  V2<T> p;
  M22<T> R;
};
template <typename T>
B2G_HD V2<T> xmul(XFM<T> t, V2<T> o) {
  return v2<T>(t.p.x + t.R.c1x * o.x + t.R.c2x * o.y, t.p.y + t.R.c1y * o.x + t.R.c2y * o.y);
}
template <typename T>
B2G_HD V2<T> xmul_t(XFM<T> t, V2<T> v) { return mmul_t(t.R, v - t.p); }
template <typename T>
B2G_HD M22<T> as_m22(Rot<T> r) {
  M22<T> m;
  m.c1x = r.c;
  m.c1y = r.s;
  m.c2x = -r.s;
  m.c2y = r.c;
  return m;
}
// Transform.mul_t(Transform) with the reference's translation (quirk Q4):
// (p_b - p_a, R_a^T R_b), the translation is not rotated.
template <typename T>
B2G_HD XFM<T> xmul_t_xf(XF<T> a, XF<T> b) {
  M22<T> A = as_m22(a.q), B = as_m22(b.q);
  XFM<T> r;
  r.p = b.p - a.p;
  r.R.c1x = A.c1x * B.c1x + A.c1y * B.c1y;
  r.R.c1y = A.c2x * B.c1x + A.c2y * B.c1y;
  r.R.c2x = A.c1x * B.c2x + A.c1y * B.c2y;
  r.R.c2y = A.c2x * B.c2x + A.c2y * B.c2y;
  return r;
}

// Rotation of an (unwrapped, FP64) angle.  FP32 pipelines reduce the angle to
// [-pi, pi] in FP64 before the FP32 sin/cos, since angles are never wrapped by
// the step (Sweep.normalize is only used by TOI, toi.py:252-253).
template <typename T>
__device__ __forceinline__ Rot<T> rot_of(double a);
template <>
__device__ __forceinline__ Rot<double> rot_of<double>(double a) {
  Rot<double> r;
  sincos(a, &r.s, &r.c);
  return r;
}
template <>
__device__ __forceinline__ Rot<float> rot_of<float>(double a) {
  const double two_pi = 6.283185307179586476925;
  double k = rint(a * (1.0 / two_pi));
  double ar = a - k * two_pi;
  Rot<float> r;
  sincosf((float)ar, &r.s, &r.c);
  return r;
}

// Python list.index(min(xs)) / list.index(max(xs)).
template <typename T>
B2G_HD int argmin_first(const T* xs, int n) {
  T m = xs[0];
  for (int i = 1; i < n; ++i) m = (xs[i] < m) ? xs[i] : m;
  for (int i = 0; i < n; ++i)
    if (xs[i] == m) return i;
  return 0;
}
template <typename T>
B2G_HD int argmax_first(const T* xs, int n) {
  T m = xs[0];
  for (int i = 1; i < n; ++i) m = (xs[i] > m) ? xs[i] : m;
  for (int i = 0; i < n; ++i)
    if (xs[i] == m) return i;
  return 0;
}
B2G_HD int pymod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// Fat-AABB overlap, PyAABB.test_overlap (common.py:830-839), in FP64.
B2G_HD bool aabb_overlap(double alx, double aly, double aux, double auy, double blx, double bly, double bux,
                         double buy) {
  if (blx - aux > 0.0 || bly - auy > 0.0) return false;
  if (alx - bux > 0.0 || aly - buy > 0.0) return false;
  return true;
}

}  // namespace b2g
